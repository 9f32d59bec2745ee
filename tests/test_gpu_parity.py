"""GPU parity of the HIP path against the FP64 oracle, through the C ABI.

The GPU evaluates integrals with spherical Hermite bra matrices, Hobson-
transformed auxiliary functions and closed-form R_tuv; the oracle with
Cartesian McMurchie-Davidson on every centre and a final spherical transform.
Tolerances are written per test.
"""

import ctypes
import functools

import numpy as np
import pytest

import oracle as O
from pysisyphus_addons_amd import _native as N
from pysisyphus_addons_amd import synthetic
from pysisyphus_addons_amd.diabatization import _intor, intor
from pysisyphus_addons_amd.intor import intor as sintor
from pysisyphus_addons_amd.shells import Shell, Shells

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def dense3c(shells, aux, screening=N.SCREEN_SCHWARZ):
    b = N.BasisArrays(*shells.as_arrays())
    x = N.BasisArrays(*aux.as_arrays())
    out = np.zeros((shells.nbf, shells.nbf, aux.nbf))
    opt = N.default_options(screening=screening)
    st = N.Stats()
    N.check(N.lib().dfhip_int3c2e_dense(ctypes.byref(b.struct), ctypes.byref(x.struct),
                                        ctypes.byref(opt), out.ctypes.data, ctypes.byref(st)))
    return out, st.as_dict()


def oracle_dense3c(shells, aux, screened=True, qvl=False):
    bas, x = O.Basis.from_shells(shells), O.Basis.from_shells(aux)
    qb, qa = O.schwarz_bra(bas), O.schwarz_aux(x)
    if qvl:
        pr, xr = O.qvl_records(bas, x, qb, qa)
    out = np.zeros((shells.nbf, shells.nbf, aux.nbf))
    for a in range(bas.nshells):
        for b in range(a + 1):
            k = a * (a + 1) // 2 + b
            for c in range(x.nshells):
                if qvl:
                    est = O.lib().orc_qvl_estimate(np.ascontiguousarray(pr[k]), np.ascontiguousarray(xr[c]),
                                                   np.ascontiguousarray(aux.shells[c].center), aux.shells[c].L)
                    if est < O.SCHWARZ_THRESH:
                        continue
                elif screened and qb[k] * qa[c] < O.SCHWARZ_THRESH:
                    continue
                blk = O.int3c2e_block(bas, x, a, b, c)
                sa = slice(bas.off[a], bas.off[a] + blk.shape[0])
                sb = slice(bas.off[b], bas.off[b] + blk.shape[1])
                sc = slice(x.off[c], x.off[c] + blk.shape[2])
                out[sa, sb, sc] = blk
                out[sb, sa, sc] = blk.transpose(1, 0, 2)
    return out


def test_boys_probe():
    """Boys F_m(T), m <= 12, T log-uniform in [1e-6, 120] (the Boys probe of
    the 60-atom workload), plus orders to 17 and the table/asymptotic seam."""
    rng = np.random.default_rng(20260001)
    n = rng.integers(0, 13, 200000).astype(np.int32)
    T = np.exp(rng.uniform(np.log(1e-6), np.log(120.0), n.size))
    extra_T = np.array([0.0, 0.025, 39.97, 39.99, 40.0, 40.01, 60.0, 1e-14])
    T = np.concatenate([T, np.repeat(extra_T, 18)])
    n = np.concatenate([n, np.tile(np.arange(18, dtype=np.int32), extra_T.size)])
    F = np.zeros_like(T)
    N.check(N.lib().dfhip_boys(T.size, n.ctypes.data, T.ctypes.data, F.ctypes.data))
    ref = np.array([O.boys(int(k), np.array([t]))[0, k] for k, t in zip(n[-200:], T[-200:])])
    Fref = O.boys(22, T)[np.arange(T.size), n]
    np.testing.assert_allclose(F[-200:], ref, rtol=1e-14, atol=0)
    err = np.abs(F - Fref)
    _say(f"boys: max abs {err.max():.2e}, max rel {(err / Fref).max():.2e}")
    assert err.max() <= 1e-14
    assert (err / Fref).max() <= 5e-14


def _mixed_system(seed=3):
    rng = np.random.default_rng(seed)
    orb, aux = [], []
    centers = rng.uniform(-1.5, 1.5, (3, 3))
    for ci, C in enumerate(centers):
        for L in (0, 1, 2, 3):
            e = 0.3 * 3.0 ** np.arange(2) * rng.uniform(0.5, 2.0)
            orb.append(Shell(L=L, center=C, exps=e, coeffs=rng.uniform(0.2, 1.0, 2), center_ind=ci))
        for L in (0, 1, 2, 3, 4, 5):
            aux.append(Shell(L=L, center=C, exps=[rng.uniform(0.2, 3.0)], coeffs=[1.0], center_ind=ci))
    return Shells(orb), Shells(aux)


def test_int3c2e_blocks_all_classes_unscreened():
    """Every (la, lb, lc) class with la, lb <= f and lc <= h, no screening."""
    shells, aux = _mixed_system()
    got, st = dense3c(shells, aux, screening=N.SCREEN_NONE)
    ref = oracle_dense3c(shells, aux, screened=False)
    scale = np.abs(ref).max()
    err = np.abs(got - ref).max()
    _say(f"3c2e all classes: max abs err {err:.2e} (max |int| {scale:.2e})")
    assert st["ncomputed"] == st["ntriplets"]
    assert err <= 1e-12 * max(scale, 1.0)


def test_int3c2e_blocks_g_orbitals():
    """Classes with a g orbital shell (the matrix-core kernels), aux up to h,
    unscreened."""
    rng = np.random.default_rng(17)
    orb, aux = [], []
    for ci, C in enumerate(rng.uniform(-1.2, 1.2, (2, 3))):
        for L in (0, 2, 4):
            e = 0.4 * 3.0 ** np.arange(2) * rng.uniform(0.5, 2.0)
            orb.append(Shell(L=L, center=C, exps=e, coeffs=rng.uniform(0.2, 1.0, 2), center_ind=ci))
        for L in (0, 3, 4, 5):
            aux.append(Shell(L=L, center=C, exps=[rng.uniform(0.3, 2.0)], coeffs=[1.0], center_ind=ci))
    shells, auxs = Shells(orb), Shells(aux)
    got, st = dense3c(shells, auxs, screening=N.SCREEN_NONE)
    ref = oracle_dense3c(shells, auxs, screened=False)
    scale = np.abs(ref).max()
    err = np.abs(got - ref).max()
    _say(f"3c2e g orbitals: max abs err {err:.2e} (max |int| {scale:.2e})")
    assert err <= 1e-12 * max(scale, 1.0)


def test_int3c2e_matrix_core_contracted():
    """Matrix-core classes (f and g orbital shells, g aux) with contracted
    aux shells and orbital shells of 4 primitives (16 primitive pairs): the
    kernels' per-iteration G path, against the oracle."""
    rng = np.random.default_rng(23)
    orb, aux = [], []
    for ci, C in enumerate(rng.uniform(-1.0, 1.0, (2, 3))):
        for L, np_ in ((0, 4), (3, 2), (4, 4)):
            e = np.sort(rng.uniform(0.2, 6.0, np_))[::-1]
            orb.append(Shell(L=L, center=C, exps=e, coeffs=rng.uniform(0.2, 1.0, np_), center_ind=ci))
        for L, np_ in ((1, 2), (4, 3)):
            e = np.sort(rng.uniform(0.3, 4.0, np_))[::-1]
            aux.append(Shell(L=L, center=C, exps=e, coeffs=rng.uniform(0.2, 1.0, np_), center_ind=ci))
    shells, auxs = Shells(orb), Shells(aux)
    got, _ = dense3c(shells, auxs, screening=N.SCREEN_NONE)
    ref = oracle_dense3c(shells, auxs, screened=False)
    scale = np.abs(ref).max()
    err = np.abs(got - ref).max()
    _say(f"3c2e matrix-core, contracted: max abs err {err:.2e} (max |int| {scale:.2e})")
    assert err <= 1e-12 * max(scale, 1.0)


def test_matrix_core_vs_lds_kernels(monkeypatch):
    """The matrix-core class kernels and the LDS kernels agree on every class
    both run (all classes up to (ff|h))."""
    shells, aux = _mixed_system(8)
    a, _ = dense3c(shells, aux, screening=N.SCREEN_NONE)
    monkeypatch.setenv("DFHIP_NO_MM", "1")
    b, _ = dense3c(shells, aux, screening=N.SCREEN_NONE)
    monkeypatch.delenv("DFHIP_NO_MM")
    err = np.abs(a - b).max()
    _say(f"matrix-core vs LDS kernels: max abs diff {err:.2e}")
    assert err <= 1e-12 * max(np.abs(a).max(), 1.0)


def test_int3c2e_blocks_contracted_aux():
    """Contracted auxiliary shells (aux = orbital basis, as in the reference's
    Coulomb-tensor test) with variable primitive counts."""
    rng = np.random.default_rng(9)
    sh = []
    for ci, C in enumerate(rng.uniform(-1.0, 1.0, (2, 3))):
        for L, np_ in ((0, 3), (1, 2), (2, 1), (0, 6), (2, 4)):
            sh.append(Shell(L=L, center=C, exps=np.sort(rng.uniform(0.1, 8.0, np_))[::-1],
                            coeffs=rng.uniform(0.1, 1.0, np_), center_ind=ci))
    shells = Shells(sh)
    got, _ = dense3c(shells, shells, screening=N.SCREEN_NONE)
    ref = oracle_dense3c(shells, shells, screened=False)
    err = np.abs(got - ref).max()
    _say(f"3c2e contracted aux: max abs err {err:.2e}")
    assert err <= 1e-12 * max(np.abs(ref).max(), 1.0)


def test_metric_and_schwarz():
    shells, aux = _mixed_system(4)
    x = N.BasisArrays(*aux.as_arrays())
    M = np.zeros((aux.nbf, aux.nbf))
    N.check(N.lib().dfhip_metric(ctypes.byref(x.struct), M.ctypes.data))
    Mref = O.metric(O.Basis.from_shells(aux))
    err = np.abs(M - Mref).max()
    _say(f"metric: max abs err {err:.2e}")
    assert err <= 1e-12 * np.abs(Mref).max()
    qa = sintor.int_schwarz_aux(aux)
    np.testing.assert_allclose(qa, O.schwarz_aux(O.Basis.from_shells(aux)), rtol=1e-13)
    qb = sintor.int_schwarz_bra(shells)
    np.testing.assert_allclose(qb, O.schwarz_bra(O.Basis.from_shells(shells)), rtol=1e-12)


def test_schwarz_bra_g_shells():
    """(ab|ab) up to (gg|gg), including the chunked classes."""
    rng = np.random.default_rng(12)
    sh = [Shell(L=L, center=rng.uniform(-1, 1, 3), exps=[0.7, 2.1], coeffs=[0.5, 0.6])
          for L in (4, 3, 4, 0, 2)]
    shells = Shells(sh)
    qb = sintor.int_schwarz_bra(shells)
    np.testing.assert_allclose(qb, O.schwarz_bra(O.Basis.from_shells(shells)), rtol=1e-12)


def _schwarz_systems():
    rng = np.random.default_rng(21)
    mixed_npp = []  # 1-, 2- and 3-primitive shells up to g in one class each
    for ci, C in enumerate(rng.uniform(-1.2, 1.2, (3, 3))):
        for L in (0, 1, 2, 3, 4):
            k = 1 + (ci + L) % 3
            e = rng.uniform(0.2, 0.6) * 3.0 ** np.arange(k)
            mixed_npp.append(Shell(L=L, center=C, exps=e, coeffs=rng.uniform(0.2, 1.0, k), center_ind=ci))
    wide = [Shell(L=L, center=rng.uniform(-1, 1, 3), exps=[0.3, 0.9, 2.7, 8.1], coeffs=[0.4, 0.5, 0.6, 0.7])
            for L in (0, 1, 2)]  # 16 primitive pairs: above the matrix-core kernel's limit
    return [("mixed", _mixed_system()[0]), ("workload0", synthetic.workload(0)[0]),
            ("mixed_npp", Shells(mixed_npp)), ("wide", Shells(wide))]


@pytest.mark.parametrize("kernel", [3, 2, 1, 0])
def test_schwarz_kernel_variants(kernel, monkeypatch):
    """Matrix-core kernels with LDS-resident (3) and streamed (2) tables,
    cached (1) and chunked (0) VALU kernels against the oracle's explicit
    (ab|ab) blocks."""
    monkeypatch.setenv("DFHIP_SCHWARZ_KERNEL", str(kernel))
    for name, shells in _schwarz_systems():
        qb = sintor.int_schwarz_bra(shells)
        ref = O.schwarz_bra(O.Basis.from_shells(shells))
        rel = np.abs(qb - ref) / np.maximum(ref, 1e-300)
        _say(f"schwarz kernel {kernel} {name}: max rel err {rel.max():.2e}")
        np.testing.assert_allclose(qb, ref, rtol=1e-12)


def test_workload0_df_and_j_vs_oracle():
    shells, aux, dens = synthetic.workload(0)
    _say(f"workload 0: nbf {shells.nbf} naux {aux.nbf} ndens {dens.shape[0]}")
    b = N.BasisArrays(*shells.as_arrays())
    x = N.BasisArrays(*aux.as_arrays())
    df, J, st = _intor.contract(b, x, dens, nstates=3, want_df=True, want_j=True)
    Jref, info = O.contract_coulomb_densities_4d(shells, aux, dens, return_info=True)
    assert st["ncomputed"] == info["ncomputed"]
    assert st["nskipped"] == info["nskipped"]
    # d before the metric solve has no conditioning amplification
    errJ = np.abs(J - Jref).max()
    _say(f"workload 0: max |dJ| {errJ:.2e}, |J|max {np.abs(Jref).max():.3e}")
    assert errJ <= 1e-10
    np.testing.assert_allclose(df, info["B"], atol=1e-6 * np.abs(info["B"]).max())


@pytest.mark.parametrize("nstates", [12])
def test_many_densities_vs_oracle(nstates):
    """More than 64 densities (78 for 12 states): the tiled-transpose packing
    of C-ordered densities and the widest contraction tiles, against the
    oracle, with the densities in both layouts."""
    shells, aux, _ = synthetic.workload(0)
    rng = np.random.Generator(np.random.PCG64(77))
    dens = synthetic.random_densities(rng, nstates, shells.nbf)
    b = N.BasisArrays(*shells.as_arrays())
    x = N.BasisArrays(*aux.as_arrays())
    _, J, st = _intor.contract(b, x, dens, nstates=nstates, want_df=False, want_j=True)
    Jref, info = O.contract_coulomb_densities_4d(shells, aux, dens, return_info=True)
    errJ = np.abs(J - Jref).max()
    _say(f"workload 0, {dens.shape[0]} densities: max |dJ| {errJ:.2e}, |J|max {np.abs(Jref).max():.3e}")
    assert st["ncomputed"] == info["ncomputed"]
    assert errJ <= 1e-10
    # the reference's Fortran layout (nbf, nbf, ndens) through the drop-in entry point
    J4 = intor.contract_coulomb_densities_4d(shells, aux, dens)
    assert np.abs(J4 - Jref).max() <= 1e-10


def test_qvl_workload0_vs_oracle():
    """QVL screening (the screener's qvl_estimate) through the drop-in path:
    same kept/skipped triplets as the oracle and J at 1e-10."""
    shells, aux, dens = synthetic.workload(0)
    b = N.BasisArrays(*shells.as_arrays())
    x = N.BasisArrays(*aux.as_arrays())
    opt = N.default_options(screening=N.SCREEN_QVL)
    _, J, st = _intor.contract(b, x, dens, nstates=3, want_df=False, want_j=True, options=opt)
    Jref, info = O.contract_coulomb_densities_4d(shells, aux, dens, return_info=True, screening="qvl")
    errJ = np.abs(J - Jref).max()
    _say(f"workload 0 QVL: max |dJ| {errJ:.2e}, computed {st['ncomputed']} / {st['ntriplets']}")
    assert st["ncomputed"] == info["ncomputed"]
    assert st["nskipped"] == info["nskipped"]
    assert errJ <= 1e-10


def test_qvl_screened_blocks():
    """Every class up to (ff|h) with QVL screening: the kept blocks match the
    oracle, the screened ones are zero."""
    shells, aux = _mixed_system(seed=5)
    # spread the centres so that the far-field branch of the estimate is taken
    sh = Shells([Shell(L=s.L, center=s.center * 8.0, exps=s.exps * 0.5, coeffs=s.coeffs, center_ind=s.center_ind)
                 for s in shells.shells])
    ax = Shells([Shell(L=s.L, center=s.center * 8.0, exps=s.exps, coeffs=s.coeffs, center_ind=s.center_ind)
                 for s in aux.shells])
    got, st = dense3c(sh, ax, screening=N.SCREEN_QVL)
    ref = oracle_dense3c(sh, ax, qvl=True)
    err = np.abs(got - ref).max()
    _say(f"QVL screened blocks: max abs err {err:.2e}, computed {st['ncomputed']} / {st['ntriplets']}")
    assert 0 < st["ncomputed"] < st["ntriplets"]
    assert err <= 1e-12 * max(np.abs(ref).max(), 1.0)


def test_workload0_screened_blocks():
    shells, aux, _ = synthetic.workload(0)
    got, st = dense3c(shells, aux)
    ref = oracle_dense3c(shells, aux, screened=True)
    err = np.abs(got - ref).max()
    _say(f"workload 0 screened (ab|P): max abs err {err:.2e}")
    assert err <= 1e-12 * np.abs(ref).max()


def test_bena2_golden_through_dropin(bena2):
    shells = Shells.from_arrays(bena2["bas_centers"], bena2["bas_spec"], bena2["bas_data"])
    J = intor.contract_coulomb_densities_4d(shells, shells, bena2["densities"])
    err = np.abs(J - bena2["J_ref"]).max()
    _say(f"bena2 golden: max |dJ| {err:.2e}")
    np.testing.assert_allclose(J, bena2["J_ref"], atol=1e-10)


def test_2d_entry_matches_4d(bena2):
    shells = Shells.from_arrays(bena2["bas_centers"], bena2["bas_spec"], bena2["bas_data"])
    B = intor.contract_coulomb_densities_2d(shells, shells, bena2["densities"])
    assert B.shape == (shells.nbf, 3) and B.flags.f_contiguous
    J = intor.contract_coulomb_densities_4d(shells, shells, bena2["densities"])
    assert J[0, 1, 0, 1] == pytest.approx(B[:, 1] @ B[:, 1], rel=1e-12)


@functools.lru_cache(maxsize=None)
def _oracle_j(wid, nstates):
    shells, aux, dens = synthetic.workload(wid)
    return O.contract_coulomb_densities_4d(shells, aux, dens, return_info=True)


def test_workload1_vs_oracle():
    shells, aux, dens = synthetic.workload(1)
    _say(f"workload 1: nbf {shells.nbf} naux {aux.nbf} ndens {dens.shape[0]}")
    b = N.BasisArrays(*shells.as_arrays())
    x = N.BasisArrays(*aux.as_arrays())
    _, J, st = _intor.contract(b, x, dens, nstates=8, want_df=False, want_j=True)
    Jref, info = _oracle_j(1, 8)
    assert st["ncomputed"] == info["ncomputed"]
    errJ = np.abs(J - Jref).max()
    _say(f"workload 1: max |dJ| {errJ:.2e} |J|max {np.abs(Jref).max():.3e} "
         f"triplets {st['ncomputed']} / {st['ntriplets']}")
    assert errJ <= 1e-10


@pytest.mark.parametrize("wid,nstates", [(0, 3), (1, 8)])
def test_mixed_precision_j(wid, nstates, monkeypatch):
    """Mixed precision (FP32 integral arithmetic below the Schwarz product
    1e-3): Coulomb-tensor elements within 1e-8 Eh of the FP64 oracle.  These
    workloads have generated classes only, which keep one FP64 window unless
    DFHIP_GEN32=1 (set here) or the tuning table asks for their FP32 kernels;
    without it the mixed run equals the FP64 one."""
    shells, aux, dens = synthetic.workload(wid)
    if wid == 1:
        b = N.BasisArrays(*shells.as_arrays())
        x = N.BasisArrays(*aux.as_arrays())
        _, J0, st0 = _intor.contract(b, x, dens, nstates=nstates, want_df=False, want_j=True,
                                     options=N.default_options(precision=N.MIXED))
        _, J64, _ = _intor.contract(b, x, dens, nstates=nstates, want_df=False, want_j=True)
        assert st0["ncomputed_fp32"] == 0 and np.array_equal(J0, J64)
    monkeypatch.setenv("DFHIP_GEN32", "1")
    b = N.BasisArrays(*shells.as_arrays())
    x = N.BasisArrays(*aux.as_arrays())
    opt = N.default_options(precision=N.MIXED)
    _, J, st = _intor.contract(b, x, dens, nstates=nstates, want_df=False, want_j=True, options=opt)
    Jref, info = _oracle_j(wid, nstates)
    errJ = np.abs(J - Jref).max()
    _say(f"workload {wid} mixed: max |dJ| {errJ:.2e} max rel {(np.abs(J - Jref) / np.abs(Jref).max()).max():.2e} "
         f"FP32 triplets {st['ncomputed_fp32']} / {st['ncomputed']}")
    assert st["ncomputed"] == info["ncomputed"]
    assert st["ncomputed_fp32"] > 0
    assert errJ <= 1e-8


def test_mixed_precision_blocks(monkeypatch):
    """(ab|P) blocks of the mixed path with the FP32 windows of the generated
    kernels (DFHIP_GEN32=1): FP64 blocks exact, FP32 blocks (all with
    Schwarz product < 1e-3) within 2e-9 absolute."""
    monkeypatch.setenv("DFHIP_GEN32", "1")
    shells, aux, _ = synthetic.workload(0)
    b = N.BasisArrays(*shells.as_arrays())
    x = N.BasisArrays(*aux.as_arrays())
    got = np.zeros((shells.nbf, shells.nbf, aux.nbf))
    opt = N.default_options(precision=N.MIXED)
    st = N.Stats()
    N.check(N.lib().dfhip_int3c2e_dense(ctypes.byref(b.struct), ctypes.byref(x.struct), ctypes.byref(opt),
                                        got.ctypes.data, ctypes.byref(st)))
    ref = oracle_dense3c(shells, aux)
    err = np.abs(got - ref)
    nz = np.abs(ref) > 0
    _say(f"mixed blocks: max abs err {err.max():.2e}, max rel err {(err[nz] / np.abs(ref[nz])).max():.2e}, "
         f"FP32 triplets {st.ncomputed_fp32} / {st.ncomputed}")
    assert st.ncomputed_fp32 > 0
    assert err.max() <= 2e-9


def test_batch_matches_single_calls():
    """The batched form (molecules on concurrent lanes) reproduces one call per
    molecule bitwise; two molecules against the oracle; the mixed batch within
    1e-8 of FP64."""
    mols, raw = [], []
    for i in range(7):
        sh, ax, dens = synthetic.batch_molecule(i)
        raw.append((sh, ax, dens))
        mols.append((N.BasisArrays(*sh.as_arrays()), N.BasisArrays(*ax.as_arrays()), dens))
    Js, sts = _intor.contract_batch(mols, [4] * len(mols), nlanes=3)
    for (b, x, dens), J, st in zip(mols, Js, sts):
        _, J1, st1 = _intor.contract(b, x, dens, nstates=4, want_df=False, want_j=True)
        assert np.array_equal(J, J1)
        assert st["ncomputed"] == st1["ncomputed"]
    small = sorted(range(len(raw)), key=lambda k: raw[k][0].nbf)[:2]
    for k in small:
        Jref = O.contract_coulomb_densities_4d(*raw[k])
        err = np.abs(Js[k] - Jref).max()
        _say(f"batch molecule {k}: nbf {raw[k][0].nbf}, max |dJ| vs oracle {err:.2e}")
        assert err <= 1e-10
    Jm, _ = _intor.contract_batch(mols, [4] * len(mols), options=N.default_options(precision=N.MIXED))
    errm = max(np.abs(a - b).max() for a, b in zip(Jm, Js))
    _say(f"mixed batch: max |dJ| vs FP64 {errm:.2e}")
    assert errm <= 1e-8


def test_column_blocks_and_device_densities():
    """The staging matrix in several column blocks (a small max_tile_bytes)
    and densities already in HBM (dens_on_device, both layouts) give the
    one-block host-density result and the oracle's J."""
    hip = ctypes.CDLL("libamdhip64.so")  # device buffers straight from the HIP runtime
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]

    shells, aux, dens = synthetic.workload(0)
    b = N.BasisArrays(*shells.as_arrays())
    x = N.BasisArrays(*aux.as_arrays())
    _, J0, st0 = _intor.contract(b, x, dens, nstates=3, want_df=False, want_j=True)
    Jref = O.contract_coulomb_densities_4d(shells, aux, dens)
    assert np.abs(J0 - Jref).max() <= 1e-10
    # about five column blocks
    tile = st0["npair_functions"] * 8 * max(1, aux.nbf // 5)
    _, J1, st1 = _intor.contract(b, x, dens, nstates=3, want_df=False, want_j=True,
                                 options=N.default_options(max_tile_bytes=tile))
    _say(f"column blocks: {st1['ncolumn_blocks']}, max |dJ| vs one block {np.abs(J1 - J0).max():.2e}")
    assert st1["ncolumn_blocks"] >= 3
    assert st1["ncomputed"] == st0["ncomputed"]
    assert np.abs(J1 - J0).max() <= 1e-12 * max(np.abs(J0).max(), 1.0)
    ndens, nbf = dens.shape[0], dens.shape[1]
    # the memory of (ndens, nbf, nbf) in C order and in column-major order
    for layout, host in ((N.DENS_C, np.ascontiguousarray(dens).ravel()),
                         (N.DENS_FORTRAN, np.asfortranarray(dens).ravel(order="F"))):
        d = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(d), host.nbytes) == 0
        try:
            assert hip.hipMemcpy(d, host.ctypes.data, host.nbytes, 1) == 0  # host to device
            J2 = np.zeros((3,) * 4, order="F")
            st = N.Stats()
            opt = N.default_options(dens_layout=layout, dens_on_device=1)
            N.check(N.lib().dfhip_contract_coulomb_densities(
                ctypes.byref(b.struct), ctypes.byref(x.struct), ndens, nbf, d, ctypes.byref(opt), None, 3,
                J2.ctypes.data, ctypes.byref(st)))
        finally:
            hip.hipFree(d)
        err = np.abs(J2 - J0).max()
        _say(f"device densities, layout {layout}: max |dJ| vs host densities {err:.2e}")
        assert err <= 1e-12 * max(np.abs(J0).max(), 1.0)


def test_bitwise_reproducible():
    shells, aux, dens = synthetic.workload(0)
    J1 = intor.contract_coulomb_densities_4d(shells, aux, dens)
    J2 = intor.contract_coulomb_densities_4d(shells, aux, dens)
    assert np.array_equal(J1, J2)


def test_bitwise_reproducible_g_shells():
    """g orbital shells take the multi-wave and the streamed Schwarz kernels
    ((ff|ff) and larger): their Schwarz norms, and so the pair order, the
    screened set and J, are the same bit for bit from call to call."""
    shells, aux, dens = synthetic.workload4_reduced(6)
    q1, q2 = sintor.int_schwarz_bra(shells), sintor.int_schwarz_bra(shells)
    assert np.array_equal(q1, q2)
    b = N.BasisArrays(*shells.as_arrays())
    x = N.BasisArrays(*aux.as_arrays())
    opt = N.default_options(screening=N.SCREEN_QVL, thresh=1e-10)
    _, J1, _ = _intor.contract(b, x, dens, nstates=20, want_df=False, want_j=True, options=opt)
    _, J2, _ = _intor.contract(b, x, dens, nstates=20, want_df=False, want_j=True, options=opt)
    assert np.array_equal(J1, J2)


def test_generated_vs_lds_kernels(monkeypatch):
    """The register-resident generated kernels and the LDS kernels agree on
    every class both implement (workload 0 plus an all-class system)."""
    shells, aux = _mixed_system(6)
    a, _ = dense3c(shells, aux, screening=N.SCREEN_NONE)
    monkeypatch.setenv("DFHIP_NO_GEN", "1")
    b, _ = dense3c(shells, aux, screening=N.SCREEN_NONE)
    monkeypatch.delenv("DFHIP_NO_GEN")
    err = np.abs(a - b).max()
    _say(f"generated vs LDS kernels: max abs diff {err:.2e}")
    assert err <= 1e-13 * max(np.abs(b).max(), 1.0)
