"""FP64 CPU oracle for the density-fitting Coulomb path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker and the CPU baseline.  The
product never imports it.

The integrals come from oracle.c (McMurchie-Davidson on all centres plus a
Cartesian->spherical transform, see its header).  This module restates the
remaining steps of the reference driver:

* ``inv_chol``: LAPACK dpotrf('L') then dtrtri('L','N')  (common/linalg.f90:60-82)
* ``dtrmm('L','L','N','N')``: B = L^-1 d             (diabatization/intor.f90:198)
* the 8-fold Coulomb-tensor assembly                 (diabatization/intor.f90:232-269)
* the Schwarz skip rule with the single-precision literal 1e-10
                                                     (diabatization/intor.f90:140)
"""

import ctypes
import math
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"

# diabatization/intor.f90:140 compares a double against the real(4) literal 1e-10
SCHWARZ_THRESH = float(np.float32(1e-10))

_lib = None

_I32P = np.ctypeslib.ndpointer(dtype=np.int32, flags="F_CONTIGUOUS")
_F64P = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_F64F = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.orc_basis_new.restype = vp
        L.orc_basis_new.argtypes = [ci, _I32P, _I32P, _F64P]
        L.orc_basis_free.argtypes = [vp]
        L.orc_basis_nbf.argtypes = [vp]
        L.orc_basis_lmax.argtypes = [vp]
        L.orc_basis_shell_off.argtypes = [vp, ci]
        L.orc_basis_shell_L.argtypes = [vp, ci]
        L.orc_basis_normed_coefs.argtypes = [vp, ci, _F64P]
        L.orc_boys.argtypes = [ci, ci, _F64P, _F64P]
        L.orc_sph_coeffs.argtypes = [ci, _F64P]
        L.orc_int3c2e_block.argtypes = [vp, vp, ci, ci, ci, _F64P]
        L.orc_int2c2e_block.argtypes = [vp, ci, ci, _F64P]
        L.orc_int4c2e_block.argtypes = [vp, ci, ci, ci, ci, _F64P]
        L.orc_schwarz_bra.argtypes = [vp, _F64P]
        L.orc_schwarz_aux.argtypes = [vp, _F64P]
        L.orc_metric.argtypes = [vp, _F64F]
        L.orc_overlap.argtypes = [vp, _F64F]
        L.orc_qvl_records.argtypes = [vp, vp, _F64P, _F64P, cd, cd, _F64P, _F64P]
        L.orc_qvl_estimate.restype = cd
        L.orc_qvl_estimate.argtypes = [_F64P, _F64P, _F64P, ci]
        L.orc_df_contract.argtypes = [vp, vp, _F64P, _F64P, vp, vp, cd, ci, _F64F, _F64F, ci, ci, ci,
                                      ctypes.POINTER(ctypes.c_int64),
                                      ctypes.POINTER(ctypes.c_int64)]
        L.orc_erfinv.restype = cd
        L.orc_erfinv.argtypes = [cd]
        L.orc_contracted_center.argtypes = [ci, _F64P, _F64P, _F64P, ci, _F64P, _F64P, _F64P,
                                            _F64P]
        L.orc_contracted_extent.restype = cd
        L.orc_contracted_extent.argtypes = [ci, _F64P, _F64P, _F64P, ci, _F64P, _F64P, _F64P, cd]
        L.orc_auxiliary_extent.restype = cd
        L.orc_auxiliary_extent.argtypes = [ci, _F64P, cd]
        L.orc_prim_cart_eri.argtypes = [ci, ci, ci, ci, cd, cd, cd, cd, _F64P, _F64P, _F64P,
                                        _F64P, _F64P]
        _lib = L
    return _lib


def ncart(L):
    return (L + 1) * (L + 2) // 2


class Basis:
    """Oracle-side shell table built from the reference basis arrays."""

    def __init__(self, bas_centers, bas_spec, bas_data):
        self.bas_centers = np.asfortranarray(bas_centers, dtype=np.int32)
        self.bas_spec = np.asfortranarray(bas_spec, dtype=np.int32)
        self.bas_data = np.ascontiguousarray(bas_data, dtype=np.float64)
        self.nshells = len(self.bas_centers)
        self.h = lib().orc_basis_new(self.nshells, self.bas_centers, self.bas_spec, self.bas_data)
        self.nbf = lib().orc_basis_nbf(self.h)
        self.lmax = lib().orc_basis_lmax(self.h)
        self.L = np.array([lib().orc_basis_shell_L(self.h, i) for i in range(self.nshells)])
        self.off = np.array([lib().orc_basis_shell_off(self.h, i) for i in range(self.nshells)])

    @classmethod
    def from_shells(cls, shells):
        return cls(*shells.as_arrays(fortran=True))

    def __del__(self):
        try:
            lib().orc_basis_free(self.h)
        except Exception:
            pass


def boys(nmax, T):
    T = np.ascontiguousarray(T, dtype=np.float64).ravel()
    F = np.zeros((T.size, nmax + 1))
    lib().orc_boys(nmax, T.size, T, F)
    return F


def sph_coeffs(L):
    out = np.zeros((2 * L + 1, ncart(L)))
    lib().orc_sph_coeffs(L, out)
    return out


def int3c2e_block(bas, aux, a, b, c):
    na, nb, nc = 2 * bas.L[a] + 1, 2 * bas.L[b] + 1, 2 * aux.L[c] + 1
    out = np.zeros(ncart(bas.L[a]) * ncart(bas.L[b]) * ncart(aux.L[c]))
    lib().orc_int3c2e_block(bas.h, aux.h, a, b, c, out)
    return out[: na * nb * nc].reshape(na, nb, nc)


def int2c2e_block(aux, P, Q):
    n1, n2 = 2 * aux.L[P] + 1, 2 * aux.L[Q] + 1
    out = np.zeros(ncart(aux.L[P]) * ncart(aux.L[Q]))
    lib().orc_int2c2e_block(aux.h, P, Q, out)
    return out[: n1 * n2].reshape(n1, n2)


def int4c2e_block(bas, a, b, c, d):
    dims = [2 * bas.L[i] + 1 for i in (a, b, c, d)]
    out = np.zeros(int(np.prod([ncart(bas.L[i]) for i in (a, b, c, d)])))
    lib().orc_int4c2e_block(bas.h, a, b, c, d, out)
    return out[: int(np.prod(dims))].reshape(dims)


def prim_cart_eri(la, lb, lc, ld, a, b, c, d, A, B, C, D):
    out = np.zeros(ncart(la) * ncart(lb) * ncart(lc) * ncart(ld))
    f = lambda x: np.ascontiguousarray(x, dtype=np.float64)
    lib().orc_prim_cart_eri(la, lb, lc, ld, a, b, c, d, f(A), f(B), f(C), f(D), out)
    return out.reshape(ncart(la), ncart(lb), ncart(lc), ncart(ld))


def schwarz_bra(bas):
    out = np.zeros(bas.nshells * (bas.nshells + 1) // 2)
    lib().orc_schwarz_bra(bas.h, out)
    return out


def schwarz_aux(aux):
    out = np.zeros(aux.nshells)
    lib().orc_schwarz_aux(aux.h, out)
    return out


def metric(aux):
    M = np.zeros((aux.nbf, aux.nbf), order="F")
    lib().orc_metric(aux.h, M)
    return M


def overlap(bas):
    S = np.zeros((bas.nbf, bas.nbf), order="F")
    lib().orc_overlap(bas.h, S)
    return S


QVL_THRESH_WS = 1e-4  # t_df_screener(shells, shells_aux, 1d-4), diabatization/intor.f90:84
QVL_AUX_PREC = 1e-8   # auxiliary_extent(exps, 1d-8), common/screening.f90:109


def qvl_records(bas, aux, qbra, qaux):
    """Screener set-up of common/screening.f90:101-163 as QVL records: per
    packed pair k = a(a+1)/2 + b (a >= b) [centre x, y, z, contracted extent,
    (min a + min b)^(1/4), Q_ab, 0, 0]; per aux shell [auxiliary extent,
    beta = min exp^(-(2L+3)/4) sqrt((2L-1)!!), Q_C, 0]."""
    pr = np.zeros((bas.nshells * (bas.nshells + 1) // 2, 8))
    xr = np.zeros((aux.nshells, 4))
    lib().orc_qvl_records(bas.h, aux.h, np.ascontiguousarray(qbra, dtype=np.float64),
                          np.ascontiguousarray(qaux, dtype=np.float64), QVL_THRESH_WS, QVL_AUX_PREC,
                          pr, xr)
    return pr, xr


def df_contract(bas, aux, densities, qbra=None, qaux=None, thresh=SCHWARZ_THRESH,
                c_begin=0, c_end=None, c_step=1, qvl=None):
    """Screened integral-direct d(P, I) = sum (mu nu|P) D_I(mu, nu), the loop of
    diabatization/intor.f90:108-183.  densities is (ndens, nbf, nbf) in any
    memory order.  qvl = (pair records, aux records) of qvl_records switches
    the screen to the QVL estimate.  Returns (d (naux, ndens) F-order,
    nskipped, ncomputed)."""
    if qbra is None:
        qbra = schwarz_bra(bas)
    if qaux is None:
        qaux = schwarz_aux(aux)
    dens = np.asfortranarray(densities, dtype=np.float64)
    ndens = dens.shape[0]
    d = np.zeros((aux.nbf, ndens), order="F")
    ns, nc = ctypes.c_int64(0), ctypes.c_int64(0)
    if c_end is None:
        c_end = aux.nshells
    qp = qx = None
    if qvl is not None:
        qp, qx = (np.ascontiguousarray(r, dtype=np.float64) for r in qvl)
    lib().orc_df_contract(bas.h, aux.h, np.ascontiguousarray(qbra), np.ascontiguousarray(qaux),
                          None if qp is None else qp.ctypes.data,
                          None if qx is None else qx.ctypes.data,
                          thresh, ndens, dens, d, c_begin, c_end, c_step, ctypes.byref(ns),
                          ctypes.byref(nc))
    return d, ns.value, nc.value


def inv_chol(M):
    """common/linalg.f90:60-82: L^-1 in the lower triangle (upper keeps M)."""
    from scipy.linalg import lapack

    c, info = lapack.dpotrf(np.asfortranarray(M), lower=1, clean=0)
    if info != 0:
        raise np.linalg.LinAlgError(f"dpotrf info={info}")
    inv, info = lapack.dtrtri(c, lower=1, unitdiag=0)
    if info != 0:
        raise np.linalg.LinAlgError(f"dtrtri info={info}")
    return inv


def dtrmm_lower(Linv, d):
    """dtrmm('L','L','N','N', naux, ndens, 1, Linv, naux, d, naux)."""
    from scipy.linalg import blas

    return blas.dtrmm(1.0, np.asfortranarray(Linv), np.asfortranarray(d), side=0, lower=1,
                      trans_a=0, diag=0)


def contract_coulomb_densities_2d(shells, shells_aux, densities, thresh=SCHWARZ_THRESH,
                                  return_info=False, screening="schwarz"):
    bas = Basis.from_shells(shells)
    aux = Basis.from_shells(shells_aux)
    qvl = None
    if screening == "qvl":
        qvl = qvl_records(bas, aux, schwarz_bra(bas), schwarz_aux(aux))
    elif screening != "schwarz":
        raise ValueError(screening)
    d, nskip, ncomp = df_contract(bas, aux, densities, thresh=thresh, qvl=qvl)
    M = metric(aux)
    Linv = inv_chol(M)
    B = dtrmm_lower(Linv, d)
    if return_info:
        return B, dict(nskipped=nskip, ncomputed=ncomp, d=d, metric=M)
    return B


def assemble_4d(B, nstates):
    """diabatization/intor.f90:241-269, 8-fold symmetric J from packed B."""
    J = np.zeros((nstates,) * 4, order="F")
    d1 = 0
    for i in range(nstates):
        for j in range(i + 1):
            d2 = 0
            for k in range(i + 1):
                lmax = j if k == i else k
                for l in range(lmax + 1):
                    val = float(B[:, d1] @ B[:, d2])
                    for idx in ((i, j, k, l), (j, i, k, l), (i, j, l, k), (j, i, l, k),
                                (k, l, i, j), (k, l, j, i), (l, k, i, j), (l, k, j, i)):
                        J[idx] = val
                    d2 += 1
            d1 += 1
    return J


def contract_coulomb_densities_4d(shells, shells_aux, densities, thresh=SCHWARZ_THRESH,
                                  return_info=False, screening="schwarz"):
    ndens = densities.shape[0]
    nstates = int(math.isqrt(8 * ndens + 1) / 2.0 - 0.5)
    assert nstates * (nstates + 1) // 2 == ndens
    B, info = contract_coulomb_densities_2d(shells, shells_aux, densities, thresh,
                                            return_info=True, screening=screening)
    J = assemble_4d(B, nstates)
    if return_info:
        info["B"] = B
        return J, info
    return J


def dense3c(bas, aux, qbra=None, qaux=None, thresh=SCHWARZ_THRESH):
    """Schwarz-screened (mu nu|P) as a dense (nbf, nbf, naux) array, both
    (mu, nu) orders filled, screened-out blocks zero (the triplet loop and
    skip rule of std/intor.f90:202-230)."""
    if qbra is None:
        qbra = schwarz_bra(bas)
    if qaux is None:
        qaux = schwarz_aux(aux)
    out = np.zeros((bas.nbf, bas.nbf, aux.nbf))
    for a in range(bas.nshells):
        for b in range(a + 1):
            k = a * (a + 1) // 2 + b
            for c in range(aux.nshells):
                if qbra[k] * qaux[c] < thresh:
                    continue
                blk = int3c2e_block(bas, aux, a, b, c)
                sa = slice(bas.off[a], bas.off[a] + blk.shape[0])
                sb = slice(bas.off[b], bas.off[b] + blk.shape[1])
                sc = slice(aux.off[c], aux.off[c] + blk.shape[2])
                out[sa, sb, sc] = blk
                out[sb, sa, sc] = blk.transpose(1, 0, 2)
    return out


def _linv_t(aux):
    """(L^-1)^T of the metric's Cholesky factor, from inv_chol's lower triangle
    (the dtrmm('R','L','T','N') of std/intor.f90:110,274)."""
    return np.tril(inv_chol(metric(aux))).T


def df1c(shells, shells_aux, thresh=SCHWARZ_THRESH):
    """std/intor.f90:15-112: D1(mu, Q) = (mu mu|Q) from the diagonal of each
    Schwarz-screened (aa|C) block, then df1c = D1 L^-T.  Returns (df1c, D1)."""
    bas, aux = Basis.from_shells(shells), Basis.from_shells(shells_aux)
    qb, qa = schwarz_bra(bas), schwarz_aux(aux)
    D1 = np.zeros((bas.nbf, aux.nbf))
    for a in range(bas.nshells):
        k = a * (a + 1) // 2 + a
        for c in range(aux.nshells):
            if qb[k] * qa[c] < thresh:
                continue
            blk = int3c2e_block(bas, aux, a, a, c)
            n = blk.shape[0]
            D1[bas.off[a]:bas.off[a] + n, aux.off[c]:aux.off[c] + blk.shape[2]] = blk[np.arange(n), np.arange(n)]
    return D1 @ _linv_t(aux), D1


def eri2c(shells, shells_aux, thresh=SCHWARZ_THRESH):
    """std/intor.f90:114-131: df1c df1c^T."""
    d, _ = df1c(shells, shells_aux, thresh)
    return d @ d.T


def df2c_mo(shells, shells_aux, mo_tensor, thresh=SCHWARZ_THRESH):
    """std/intor.f90:133-292 without its drops of single values <= 1e-10:
    row i(i+1)/2 + a (i >= a) of sum_{mu nu} C(mu, i) C(nu, a) (mu nu|Q),
    times L^-T.  Returns (df2c_mo, the unpacked (nmo, nmo, naux) MO tensor)."""
    bas, aux = Basis.from_shells(shells), Basis.from_shells(shells_aux)
    T = dense3c(bas, aux, thresh=thresh)
    C = np.asarray(mo_tensor, dtype=np.float64)
    X = np.einsum("mnQ,mi,na->iaQ", T, C, C, optimize=True)
    nmo = C.shape[1]
    rows = np.array([X[i, a] for i in range(nmo) for a in range(i + 1)])
    return rows @ _linv_t(aux), X


def int3c2e_estimates(shells, shells_aux):
    """intor/benchmark.f90:15-132 restated: for a >= b (the reference's
    1-based loops a = 1..n, b = 1..a), pairs neither concentric (|A - B| <=
    1e-6) nor small (Q_ab <= 1e-10), every aux shell c: [Q_ab Q_C, QVL
    estimate, ||(ab|C)||_F, |contracted centre - C|] (screening.f90:315-326).
    Returns (records (n, 4), nskipped with QVL estimate < the real(4) 1e-10)."""
    bas, aux = Basis.from_shells(shells), Basis.from_shells(shells_aux)
    qb, qa = schwarz_bra(bas), schwarz_aux(aux)
    pr, xr = qvl_records(bas, aux, qb, qa)
    recs, nskip = [], 0
    for a in range(bas.nshells):
        A = np.asarray(shells.shells[a].center, dtype=float)
        for b in range(a + 1):
            B = np.asarray(shells.shells[b].center, dtype=float)
            k = a * (a + 1) // 2 + b
            if np.linalg.norm(A - B) <= 1e-6 or qb[k] <= 1e-10:
                continue
            for c in range(aux.nshells):
                C = np.ascontiguousarray(shells_aux.shells[c].center, dtype=float)
                q = lib().orc_qvl_estimate(np.ascontiguousarray(pr[k]), np.ascontiguousarray(xr[c]), C,
                                           int(aux.L[c]))
                nskip += q < SCHWARZ_THRESH
                recs.append([qb[k] * qa[c], q, np.linalg.norm(int3c2e_block(bas, aux, a, b, c)),
                             np.linalg.norm(pr[k, :3] - C)])
    return np.array(recs).reshape(-1, 4), nskip


# grid/grid.f90:9-13: primitives with -a r^2 below CUTOFF are skipped; the
# Cartesian components carry 1/sqrt((2lx-1)!! (2ly-1)!! (2lz-1)!!)
GRID_CUTOFF = -36.0
_DFACT_ODD = (1.0, 1.0, 3.0, 15.0, 105.0)


def cart_exponents(L):
    """Cartesian order of a shell (lx = L..0, then ly = L-lx..0): xx, xy, xz,
    yy, yz, zz for L = 2 (grid.f90:64-99)."""
    return [(lx, ly, L - lx - ly) for lx in range(L, -1, -1) for ly in range(L - lx, -1, -1)]


def cart_norm(lx, ly, lz):
    return 1.0 / math.sqrt(_DFACT_ODD[lx] * _DFACT_ODD[ly] * _DFACT_ODD[lz])


def grid_shells(bas_centers, bas_spec, bas_data):
    """(L, centre, exponents, normalized coefficients, first Cartesian index)
    per shell from the reference basis arrays (common/shells.f90:97-187); the
    coefficients come from the oracle's norm_cgto_coeffs."""
    bas = Basis(bas_centers, bas_spec, bas_data)
    bc, bs, bd = bas.bas_centers, bas.bas_spec, bas.bas_data
    out, c0 = [], 0
    for i in range(bas.nshells):
        L, npr, pp = int(bs[i, 1]), int(bs[i, 2]), int(bs[i, 4]) - 1
        cp = int(bc[i, 2]) - 1
        coefs = np.zeros(npr)
        lib().orc_basis_normed_coefs(bas.h, i, coefs)
        out.append((L, bd[cp:cp + 3].copy(), bd[pp + npr:pp + 2 * npr].copy(), coefs, c0))
        c0 += ncart(L)
    return out, c0


def cart_ao_values(shells, ncart_total, points):
    """chi[i, c] of the Cartesian functions at points (n, 3): cart_gto3d_rel
    (grid.f90:17-110) restated with numpy over the points."""
    points = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    chi = np.zeros((len(points), ncart_total))
    for L, A, exps, coefs, c0 in shells:
        d = points - A
        dist = (d * d).sum(axis=1)
        rad = np.zeros(len(points))
        for a, c in zip(exps, coefs):
            arg = -a * dist
            keep = arg >= GRID_CUTOFF
            rad[keep] += c * np.exp(arg[keep])
        for k, (lx, ly, lz) in enumerate(cart_exponents(L)):
            chi[:, c0 + k] = cart_norm(lx, ly, lz) * d[:, 0] ** lx * d[:, 1] ** ly * d[:, 2] ** lz * rad
    return chi


def eval_densities(bas_centers, bas_spec, bas_data, grid3d, densities, blk_size=100, thresh=1e-8,
                   return_masks=False):
    """grid/grid.f90:112-286 restated: the points in blocks of blk_size; per
    block the mean |chi_nu| (:222-229), the pairs nu <= mu whose
    |((f D_j(mu, nu)) mean_nu) mean_mu| >= thresh (f = 2 off the diagonal,
    :234-258), and rho_j(i) = sum over the kept pairs of f D_j(mu, nu)
    chi_mu(i) chi_nu(i) (:260-271).  densities (ndens, nbfs, nbfs) is read as
    element [j, mu, nu] for mu >= nu.  Returns (ndens, npoints); with
    return_masks also the per-block contrib arrays (for margin checks)."""
    shells, nc = grid_shells(bas_centers, bas_spec, bas_data)
    grid3d = np.asarray(grid3d, dtype=np.float64).reshape(-1, 3)
    D = np.asarray(densities, dtype=np.float64)
    ndens, npts = D.shape[0], len(grid3d)
    assert D.shape[1] == D.shape[2] == nc
    low = np.tril(np.ones((nc, nc), dtype=bool))            # [mu, nu] with mu >= nu
    fac = np.where(np.eye(nc, dtype=bool), 1.0, 2.0)
    FD = np.where(low[None], fac[None] * D, 0.0)             # f D_j(mu, nu), mu >= nu
    rho = np.zeros((ndens, npts))
    contribs = []
    for b0 in range(0, npts, blk_size):
        chi = cart_ao_values(shells, nc, grid3d[b0:b0 + blk_size])
        cur = len(chi)
        mean = np.zeros(nc)
        for i in range(cur):                                 # summed in point order
            mean += np.abs(chi[i])
        mean /= cur
        contrib = FD * mean[None, None, :] * mean[None, :, None]   # ((f D) mean_nu) mean_mu
        M = np.where(low[None] & (np.abs(contrib) >= thresh), FD, 0.0)
        if return_masks:
            contribs.append(np.where(low[None], np.abs(contrib), np.nan))
        # rho_j(i) = sum_{mu, nu} M_j[mu, nu] chi[i, mu] chi[i, nu]
        rho[:, b0:b0 + cur] = np.einsum("im,jmn,in->ji", chi, M, chi, optimize=True)
    if return_masks:
        return rho, contribs
    return rho
