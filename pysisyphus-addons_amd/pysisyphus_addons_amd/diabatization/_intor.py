"""Native layer of the Coulomb-tensor path; same functions, arguments and
return layouts as the reference's Cython module diabatization/_intor.pyx:55-176.

Arguments are the basis arrays returned by ``Shells.as_arrays(fortran=True)``
(int32 / float64, C-contiguous, as the typed memoryviews of the reference
require) and the densities (ndens, nbfs, nbfs) float64, C-contiguous.
Differences from the reference, all deliberate:
  * ``naux`` of the 2d entry is the number of spherical auxiliary functions
    (the reference sums 2*nprims+1, _intor.pyx:67-69);
  * the densities are handed to the library in their C order, without the
    full np.asfortranarray copy of _intor.pyx:147 (the contraction only uses
    D + D^T, so both orders give identical results);
  * failures raise DfhipError instead of aborting the process.
"""

import math

import numpy as np

from .. import _native as N


def _check_memoryview_like(arr, dtype, ndim, name):
    if not isinstance(arr, np.ndarray):
        arr = np.asarray(arr)
    if arr.dtype != dtype or arr.ndim != ndim or not arr.flags.c_contiguous:
        raise ValueError(f"{name}: expected C-contiguous {np.dtype(dtype).name} array with {ndim} dims")
    return arr


def _prep(bas_centers, bas_spec, bas_data, bas_centers_aux, bas_spec_aux, bas_data_aux, densities):
    _check_memoryview_like(bas_centers, np.int32, 2, "bas_centers")
    _check_memoryview_like(bas_spec, np.int32, 2, "bas_spec")
    _check_memoryview_like(bas_data, np.float64, 1, "bas_data")
    _check_memoryview_like(bas_centers_aux, np.int32, 2, "bas_centers_aux")
    _check_memoryview_like(bas_spec_aux, np.int32, 2, "bas_spec_aux")
    _check_memoryview_like(bas_data_aux, np.float64, 1, "bas_data_aux")
    dens = _check_memoryview_like(densities, np.float64, 3, "densities")
    b = N.BasisArrays(bas_centers, bas_spec, bas_data)
    x = N.BasisArrays(bas_centers_aux, bas_spec_aux, bas_data_aux)
    return b, x, dens


def contract(bas, aux, dens, nstates=0, want_df=True, want_j=False, options=None, stats=None):
    """Runs the full path; returns (df (naux, ndens) F-order or None,
    J (nstates,)*4 F-order or None, stats dict)."""
    ndens, nbfs = dens.shape[0], dens.shape[1]
    opt = options if options is not None else N.default_options()
    opt.dens_layout = N.DENS_C
    df = np.zeros((aux.nsph, ndens), order="F") if want_df else None
    J = np.zeros((nstates,) * 4, order="F") if want_j else None
    st = N.Stats()
    status = N.lib().dfhip_contract_coulomb_densities(
        N.ctypes.byref(bas.struct), N.ctypes.byref(aux.struct), ndens, nbfs, dens.ctypes.data,
        N.ctypes.byref(opt), df.ctypes.data if df is not None else None, nstates,
        J.ctypes.data if J is not None else None, N.ctypes.byref(st))
    N.check(status)
    return df, J, st.as_dict()


def contract_batch(mols, nstates, options=None, nlanes=0):
    """Coulomb tensors of independent molecules in one batched call.
    mols: list of (BasisArrays, BasisArrays, densities (ndens, nbf, nbf));
    nstates: per molecule.  Returns (list of J (nstates,)*4 F-order, list of
    stats dicts)."""
    n = len(mols)
    opt = options if options is not None else N.default_options()
    opt.dens_layout = N.DENS_C
    basT = (N.Basis * n)(*[m[0].struct for m in mols])
    auxT = (N.Basis * n)(*[m[1].struct for m in mols])
    dens = [np.ascontiguousarray(m[2], dtype=np.float64) for m in mols]
    nd = np.array([d.shape[0] for d in dens], dtype=np.int32)
    nb = np.array([d.shape[1] for d in dens], dtype=np.int32)
    ns = np.array(nstates, dtype=np.int32)
    Js = [np.zeros((int(k),) * 4, order="F") for k in ns]
    dptr = (N.ctypes.c_void_p * n)(*[d.ctypes.data for d in dens])
    jptr = (N.ctypes.c_void_p * n)(*[J.ctypes.data for J in Js])
    stats = (N.Stats * n)()
    status = N.lib().dfhip_contract_coulomb_densities_batch(
        n, basT, auxT, nd.ctypes.data, nb.ctypes.data, dptr, N.ctypes.byref(opt), ns.ctypes.data, jptr, stats,
        nlanes)
    N.check(status)
    return Js, [st.as_dict() for st in stats]


def contract_coulomb_densities_2d(bas_centers, bas_spec, bas_data, bas_centers_aux, bas_spec_aux,
                                  bas_data_aux, densities):
    b, x, dens = _prep(bas_centers, bas_spec, bas_data, bas_centers_aux, bas_spec_aux,
                       bas_data_aux, densities)
    df, _, _ = contract(b, x, dens, want_df=True)
    return df


def contract_coulomb_densities_4d(bas_centers, bas_spec, bas_data, bas_centers_aux, bas_spec_aux,
                                  bas_data_aux, densities):
    b, x, dens = _prep(bas_centers, bas_spec, bas_data, bas_centers_aux, bas_spec_aux,
                       bas_data_aux, densities)
    ndens = dens.shape[0]
    # "nstates * (nstates + 1) / 2 = ndens" solved for nstates (_intor.pyx:129-131)
    nstates = int(math.isqrt(8 * ndens + 1) / 2.0 - 0.5)
    assert nstates * (nstates + 1) // 2 == ndens
    _, J, _ = contract(b, x, dens, nstates=nstates, want_df=False, want_j=True)
    return J
