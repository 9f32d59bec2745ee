"""Native layer of the std/ DF tensors; same functions, arguments and return
layouts as the reference's std/_intor.pyx:65-228 (F-ordered outputs, nbfs and
nbfs_aux from the L column of bas_spec).  Failures raise DfhipError."""

import numpy as np

from .. import _native as N


def _nsph(spec):
    return int(np.sum(2 * np.asarray(spec)[:, 1] + 1))


def int_eri2c(bas_centers, bas_spec, bas_data, bas_centers_aux, bas_spec_aux, bas_data_aux):
    b = N.BasisArrays(bas_centers, bas_spec, bas_data)
    x = N.BasisArrays(bas_centers_aux, bas_spec_aux, bas_data_aux)
    nbfs = _nsph(b.spec)
    out = np.zeros((nbfs, nbfs), order="F")
    N.lib().f_intor_eri2c(*b.args(), *x.args(), nbfs, out.ctypes.data)
    N.check()
    return out


def int_df1c(bas_centers, bas_spec, bas_data, bas_centers_aux, bas_spec_aux, bas_data_aux):
    b = N.BasisArrays(bas_centers, bas_spec, bas_data)
    x = N.BasisArrays(bas_centers_aux, bas_spec_aux, bas_data_aux)
    nbfs, nbfs_aux = _nsph(b.spec), _nsph(x.spec)
    out = np.zeros((nbfs, nbfs_aux), order="F")
    N.lib().f_intor_df1c(*b.args(), *x.args(), nbfs, nbfs_aux, out.ctypes.data)
    N.check()
    return out


def int_df2c_mo(bas_centers, bas_spec, bas_data, bas_centers_aux, bas_spec_aux, bas_data_aux, mo_tensor):
    b = N.BasisArrays(bas_centers, bas_spec, bas_data)
    x = N.BasisArrays(bas_centers_aux, bas_spec_aux, bas_data_aux)
    C = np.asfortranarray(mo_tensor, dtype=np.float64)
    if C.ndim != 2:
        raise ValueError("mo_tensor must have shape (nbfs, nmos)")
    nbfs, nmos = C.shape
    nbfs_aux = _nsph(x.spec)
    out = np.zeros((nmos * (nmos + 1) // 2, nbfs_aux), order="F")
    N.lib().f_intor_df2c_mo(*b.args(), *x.args(), nbfs, nmos, C.ctypes.data, nbfs_aux, out.ctypes.data)
    N.check()
    return out
