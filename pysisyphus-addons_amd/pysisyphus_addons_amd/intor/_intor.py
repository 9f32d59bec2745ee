"""Native layer of the Schwarz-norm entry points; same functions and
return layouts as the reference's intor/_intor.pyx."""

import numpy as np

from .. import _native as N


def int_schwarz_aux(bas_centers, bas_spec, bas_data):
    b = N.BasisArrays(bas_centers, bas_spec, bas_data)
    out = np.zeros(len(b.centers))
    N.lib().f_int_schwarz_aux(*b.args(), out.ctypes.data)
    N.check()
    return out


def int_schwarz_bra(bas_centers, bas_spec, bas_data):
    b = N.BasisArrays(bas_centers, bas_spec, bas_data)
    n = len(b.centers)
    out = np.zeros(n * (n + 1) // 2)
    N.lib().f_int_schwarz_bra(*b.args(), out.ctypes.data)
    N.check()
    return out
