"""The reference's C boundary itself, called through ctypes exactly as the
reference's Cython layer calls it (diabatization/_intor.pyx:55-176 binding
diabatization/intor_wrapper.f90:12-76): column-major int32 basis arrays,
Fortran-ordered (ndens, nbfs, nbfs) host densities (the np.asfortranarray
copy of _intor.pyx:147), caller-zeroed Fortran-ordered outputs, scalars by
value, void return.  Statuses are read with dfhip_last_status().
"""

import ctypes
import math

import numpy as np
import pytest

import oracle as O
from pysisyphus_addons_amd import _native as N
from pysisyphus_addons_amd import synthetic
from pysisyphus_addons_amd.shells import Shells

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def _say(msg):
    print(msg, flush=True)


def _fargs(shells):
    """(nshells, ndata, bas_centers, bas_spec, bas_data) as the Cython layer
    hands them over: F-ordered int32 copies and the float64 data."""
    bc, bs, bd = shells.as_arrays(fortran=True)
    bc = np.asfortranarray(bc, dtype=np.int32)
    bs = np.asfortranarray(bs, dtype=np.int32)
    bd = np.ascontiguousarray(bd, dtype=np.float64)
    return (bc, bs, bd), (len(bc), bd.size, bc.ctypes.data, bs.ctypes.data, bd.ctypes.data)


def call_4d(shells, aux, dens):
    L = N.lib()
    L.f_init()
    keep_o, a_o = _fargs(shells)
    keep_x, a_x = _fargs(aux)
    fd = np.asfortranarray(dens)  # _intor.pyx:147
    ndens, nbfs = dens.shape[0], dens.shape[1]
    nstates = (math.isqrt(8 * ndens + 1) - 1) // 2
    assert nstates * (nstates + 1) // 2 == ndens
    J = np.zeros((nstates,) * 4, order="F")
    L.f_contract_coulomb_densities_4d(*a_o, *a_x, ndens, nbfs, fd.ctypes.data, nstates, J.ctypes.data)
    return J, L.dfhip_last_status()


def call_2d(shells, aux, dens, naux):
    L = N.lib()
    L.f_init()
    keep_o, a_o = _fargs(shells)
    keep_x, a_x = _fargs(aux)
    fd = np.asfortranarray(dens)
    ndens, nbfs = dens.shape[0], dens.shape[1]
    B = np.zeros((naux, ndens), order="F")
    L.f_contract_coulomb_densities_2d(*a_o, *a_x, ndens, nbfs, fd.ctypes.data, naux, B.ctypes.data)
    return B, L.dfhip_last_status()


def _b_tolerance(M, B):
    """Forward-error scale of B = L^-1 d: a triangular solve with the Cholesky
    factor amplifies relative perturbations by kappa(L) = sqrt(kappa(M));
    64 eps kappa(L) max|B| (two independent FP64 solvers on the host differ
    by about 1/10 of it, on workload 0 and on bena2)."""
    w = np.linalg.eigvalsh(M)
    kappa = w[-1] / w[0]
    return 64.0 * EPS * np.sqrt(kappa) * np.abs(B).max(), kappa


def test_f_4d_bena2_golden(bena2):
    """tests/test_coulomb_tensor/test_coulomb_tensor.py:7-32 through the C
    symbol: the golden J at 1e-10 and the oracle's J."""
    shells = Shells.from_arrays(bena2["bas_centers"], bena2["bas_spec"], bena2["bas_data"])
    J, status = call_4d(shells, shells, bena2["densities"])
    assert status == 0
    err = np.abs(J - bena2["J_ref"]).max()
    _say(f"f_contract_coulomb_densities_4d bena2: max |dJ| vs golden {err:.2e}")
    assert err <= 1e-10
    Jo = O.contract_coulomb_densities_4d(shells, shells, bena2["densities"])
    assert np.abs(J - Jo).max() <= 1e-10


def test_f_4d_workload0_vs_oracle():
    shells, aux, dens = synthetic.workload(0)
    J, status = call_4d(shells, aux, dens)
    assert status == 0
    Jo = O.contract_coulomb_densities_4d(shells, aux, dens)
    err = np.abs(J - Jo).max()
    _say(f"f_contract_coulomb_densities_4d workload 0: max |dJ| vs oracle {err:.2e}")
    assert err <= 1e-10


@pytest.mark.parametrize("case", ["workload0", "bena2"])
def test_f_2d_vs_oracle(case, bena2):
    """B = L^-1 d through the C symbol, against the oracle's LAPACK
    dpotrf/dtrtri/dtrmm, at the conditioning-derived tolerance; and the
    residual L_oracle B_gpu - d_oracle."""
    if case == "workload0":
        shells, aux, dens = synthetic.workload(0)
    else:
        shells = aux = Shells.from_arrays(bena2["bas_centers"], bena2["bas_spec"], bena2["bas_data"])
        dens = bena2["densities"]
    naux = aux.nbf
    B, status = call_2d(shells, aux, dens, naux)
    assert status == 0
    assert B.flags.f_contiguous and B.shape == (naux, dens.shape[0])
    Bo, info = O.contract_coulomb_densities_2d(shells, aux, dens, return_info=True)
    tol, kappa = _b_tolerance(info["metric"], Bo)
    err = np.abs(B - Bo).max()
    Lo = np.linalg.cholesky(info["metric"])
    res = np.abs(Lo @ B - info["d"]).max()
    _say(f"f_contract_coulomb_densities_2d {case}: cond(M) {kappa:.2e}, max |dB| {err:.2e} (tol {tol:.2e}), "
         f"max |L B - d| {res:.2e} (max |d| {np.abs(info['d']).max():.2e})")
    assert err <= tol
    assert res <= 1e-12 * np.abs(info["d"]).max()


def test_f_2d_reference_naux_is_rejected(bena2):
    """The reference's Cython passes naux = sum(2 nprims + 1)
    (diabatization/_intor.pyx:67-69); for bena2's contracted shells that is
    not the number of spherical aux functions: status DFHIP_ERR_ARG, the
    output untouched."""
    shells = Shells.from_arrays(bena2["bas_centers"], bena2["bas_spec"], bena2["bas_data"])
    spec = np.asarray(bena2["bas_spec"])
    naux_ref = int(np.sum(2 * spec[:, 2] + 1))
    assert naux_ref != shells.nbf
    B, status = call_2d(shells, shells, bena2["densities"], naux_ref)
    _say(f"reference naux {naux_ref} vs {shells.nbf} spherical functions: status {status}")
    assert status == 1
    assert not B.any()


def test_non_spd_metric_reports_status():
    """An auxiliary basis holding every shell twice has a singular metric:
    the Cholesky meets a pivot <= 0 (or a NaN behind a tiny one) and the call
    reports DFHIP_ERR_NOT_SPD (the reference ignores dpotrf's info,
    common/linalg.f90:78-81)."""
    shells, aux, dens = synthetic.workload(0)
    twice = Shells(list(aux.shells) + list(aux.shells))
    J, status = call_4d(shells, twice, dens)
    msg = N.lib().dfhip_last_error().decode()
    _say(f"duplicated aux basis: status {status} ({msg})")
    assert status == 3
    assert "positive definite" in msg
    # the next call on the same thread succeeds again
    J, status = call_4d(shells, aux, dens)
    assert status == 0
