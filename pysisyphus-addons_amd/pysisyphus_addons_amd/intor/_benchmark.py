"""Native layer of the screening benchmark; the function of the reference's
intor/_benchmark.pyx:27-67, plus the records in memory."""

import ctypes

import numpy as np

from .. import _native as N


def benchmark_int3c2e(bas_centers, bas_spec, bas_data, bas_centers_aux, bas_spec_aux, bas_data_aux):
    """Writes int_estimates.bin (4 float64 per triplet) in the working
    directory and prints the reference's log lines."""
    b = N.BasisArrays(bas_centers, bas_spec, bas_data)
    x = N.BasisArrays(bas_centers_aux, bas_spec_aux, bas_data_aux)
    N.lib().f_benchmark_int3c2e(*b.args(), *x.args())
    N.check()


def int3c2e_estimates(bas_centers, bas_spec, bas_data, bas_centers_aux, bas_spec_aux, bas_data_aux):
    """The records of benchmark_int3c2e as an (n, 4) array: Schwarz estimate,
    QVL estimate, ||(ab|C)||_F, distance of the pair centre from C."""
    b = N.BasisArrays(bas_centers, bas_spec, bas_data)
    x = N.BasisArrays(bas_centers_aux, bas_spec_aux, bas_data_aux)
    n = ctypes.c_int64(0)
    N.check(N.lib().dfhip_int3c2e_estimates(ctypes.byref(b.struct), ctypes.byref(x.struct), None, 0,
                                            ctypes.byref(n)))
    out = np.zeros((n.value, 4))
    N.check(N.lib().dfhip_int3c2e_estimates(ctypes.byref(b.struct), ctypes.byref(x.struct), out.ctypes.data,
                                            n.value, ctypes.byref(n)))
    return out


def metric_powerh(bas_centers, bas_spec, bas_data, power):
    """(P|Q)^power by the eigen route (common/linalg.f90:36-58)."""
    x = N.BasisArrays(bas_centers, bas_spec, bas_data)
    out = np.zeros((x.nsph, x.nsph), order="F")
    N.check(N.lib().dfhip_metric_powerh(ctypes.byref(x.struct), float(power), out.ctypes.data))
    return out
