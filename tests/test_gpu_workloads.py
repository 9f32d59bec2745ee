"""The benchmark workloads at their full sizes (and the 200-atom recipe at a
size the oracle covers), through the C ABI, against the FP64 oracle and
through size-independent properties:

* the 60-atom f/g workload (orbital 3s2p1d1f, aux 4s3p2d1f1g, 8 states):
  FP64 J against the oracle at 1e-10 with identical triplet counts, the
  mixed path at 1e-8;
* the 200-atom recipe (orbital 4s3p2d1f1g, aux 5s4p3d2f1g1h, 20 states,
  QVL at 1e-10) on 10 atoms: FP64 at 1e-10 with identical kept/skipped
  counts, mixed at 1e-8, also with the FP32 Et tables refused (the FP64
  fallback of the matrix-core windows);
* the 200-atom workload itself: FP64 against mixed at 1e-8, the 8-fold
  symmetry of J, the class histogram summing to the screened count;
* the QVL screener records of the device path against the oracle's.
"""

import ctypes
import functools

import numpy as np
import pytest

import oracle as O
from pysisyphus_addons_amd import _native as N
from pysisyphus_addons_amd import synthetic
from pysisyphus_addons_amd.diabatization import _intor

pytestmark = pytest.mark.gpu

# torch ships a HIP runtime of its own; it finds no device once the library's
# HIP runtime has initialised the GPU in the same process, so the 200-atom test
# (device densities drawn with torch) lets torch take the device first, at
# collection, before any test calls the library (bench.py does the same).
try:
    import torch

    if torch.cuda.is_available():
        torch.zeros(1, device="cuda:0")
except ImportError:
    torch = None


def _say(msg):
    print(msg, flush=True)


def _arrays(shells, aux):
    return N.BasisArrays(*shells.as_arrays()), N.BasisArrays(*aux.as_arrays())


@functools.lru_cache(maxsize=None)
def _oracle(key):
    if key == "w2":
        shells, aux, dens = synthetic.workload(2)
        screening = "schwarz"
    else:
        shells, aux, dens = synthetic.workload4_reduced(10)
        screening = "qvl"
    J, info = O.contract_coulomb_densities_4d(shells, aux, dens, return_info=True, screening=screening)
    return J, info


def _gpu(key, precision):
    if key == "w2":
        shells, aux, dens = synthetic.workload(2)
        opt = N.default_options(precision=precision)
    else:
        shells, aux, dens = synthetic.workload4_reduced(10)
        opt = N.default_options(precision=precision, screening=N.SCREEN_QVL, thresh=1e-10)
    b, x = _arrays(shells, aux)
    nstates = {"w2": 8, "w4r": 20}[key]
    _, J, st = _intor.contract(b, x, dens, nstates=nstates, want_df=False, want_j=True, options=opt)
    return J, st


def _check_counts(st):
    assert st["ncomputed"] + st["nskipped"] == st["ntriplets"]
    assert sum(st["class_triplets"]) == st["ncomputed"]


@pytest.mark.parametrize("key", ["w2", "w4r"])
def test_fp64_vs_oracle(key):
    _say(f"{key}: oracle ...")
    Jo, info = _oracle(key)
    J, st = _gpu(key, N.FP64)
    err = np.abs(J - Jo).max()
    _say(f"{key} FP64: max |dJ| {err:.2e} (|J|max {np.abs(Jo).max():.3e}), computed {st['ncomputed']} "
         f"(oracle {info['ncomputed']}), skipped {st['nskipped']} (oracle {info['nskipped']})")
    assert st["ncomputed"] == info["ncomputed"]
    assert st["nskipped"] == info["nskipped"]
    _check_counts(st)
    assert err <= 1e-10


@pytest.mark.parametrize("key", ["w2", "w4r"])
def test_mixed_vs_oracle(key):
    Jo, info = _oracle(key)
    J, st = _gpu(key, N.MIXED)
    err = np.abs(J - Jo)
    _say(f"{key} mixed: max |dJ| {err.max():.2e}, max rel {(err.max() / np.abs(Jo).max()):.2e}, "
         f"low-precision triplets {st['ncomputed_fp32']} / {st['ncomputed']}")
    assert st["ncomputed"] == info["ncomputed"]
    _check_counts(st)
    assert 0 < st["ncomputed_fp32"] <= st["ncomputed"]
    assert err.max() <= 1e-8


def test_mixed_fp32_table_fallback(monkeypatch):
    """Mixed path on the 200-atom recipe (10 atoms, QVL) with the FP32 Et
    table budget at one byte (DFHIP_ET32_CAP): every matrix-core class falls
    back to a single FP64 window, whose Et table is then built again for the
    whole prefixes (it first held only the pairs with an FP64 window); J
    against the oracle at 1e-8 with identical counts, and fewer FP32
    triplets than the default run."""
    Jo, info = _oracle("w4r")
    J0, st0 = _gpu("w4r", N.MIXED)
    monkeypatch.setenv("DFHIP_ET32_CAP", "1")
    J, st = _gpu("w4r", N.MIXED)
    err = np.abs(J - Jo).max()
    _say(f"w4r mixed, FP32 tables refused: max |dJ| {err:.2e}, low-precision triplets {st['ncomputed_fp32']} "
         f"(default run {st0['ncomputed_fp32']}) / {st['ncomputed']}")
    assert st["ncomputed"] == info["ncomputed"]
    _check_counts(st)
    assert st["ncomputed_fp32"] < st0["ncomputed_fp32"]
    assert err <= 1e-8


@pytest.mark.parametrize("key", ["w1", "w4r40"])
def test_pair_drop_changes_nothing(key, monkeypatch):
    """The pairs dropped by their Schwarz bound before the (ab|ab) blocks
    (dfhip_schwarz_bounds) keep no triplet: J and every count equal those of
    the run that evaluates every pair, bit for bit.  w1: the 60-atom s-d
    workload, Schwarz; w4r40: the 200-atom recipe on 40 atoms, QVL; both
    FP64 and mixed.  Most pairs of either are dropped."""
    if key == "w1":
        shells, aux, dens = synthetic.workload(1)
        nstates, kw = 8, {}
    else:
        shells, aux, dens = synthetic.workload4_reduced(40)
        nstates, kw = 20, dict(screening=N.SCREEN_QVL, thresh=1e-10)
    b, x = _arrays(shells, aux)
    npairs = len(shells.shells) * (len(shells.shells) + 1) // 2
    for prec in (N.FP64, N.MIXED):
        opt = N.default_options(precision=prec, **kw)
        monkeypatch.delenv("DFHIP_NO_PAIR_SKIP", raising=False)
        _, J, st = _intor.contract(b, x, dens, nstates=nstates, want_df=False, want_j=True, options=opt)
        monkeypatch.setenv("DFHIP_NO_PAIR_SKIP", "1")
        _, J0, st0 = _intor.contract(b, x, dens, nstates=nstates, want_df=False, want_j=True, options=opt)
        _say(f"{key} precision {prec}: pairs with Q_ab = 0: {st['npairs_zero_q']} dropped run, "
             f"{st0['npairs_zero_q']} full run, of {npairs}; Schwarz {st['ms_schwarz']:.2f} ms vs "
             f"{st0['ms_schwarz']:.2f} ms; max |dJ| {np.abs(J - J0).max():.1e}")
        assert st["npairs_zero_q"] > npairs // 2 > st0["npairs_zero_q"]
        for k in ("ncomputed", "nskipped", "ncomputed_fp32", "class_triplets", "class_prim_triplets"):
            assert st[k] == st0[k], k
        assert np.array_equal(J, J0)


def test_mixed_blocks_fg_classes():
    """(ab|P) blocks of the mixed path on the f/g-orbital, g/h-aux classes (the
    200-atom recipe on 4 atoms), against the oracle: max abs and rel error.
    An FP32 block has |(ab|P)| <= Q_ab Q_C < fp32_below; its error stays
    within 2e-6 of that bound (about 30 FP32 roundings)."""
    shells, aux, _ = synthetic.workload4_reduced(4)
    b, x = _arrays(shells, aux)
    got = np.zeros((shells.nbf, shells.nbf, aux.nbf))
    opt = N.default_options(precision=N.MIXED, screening=N.SCREEN_QVL, thresh=1e-10)
    st = N.Stats()
    N.check(N.lib().dfhip_int3c2e_dense(ctypes.byref(b.struct), ctypes.byref(x.struct), ctypes.byref(opt),
                                        got.ctypes.data, ctypes.byref(st)))
    from test_gpu_parity import oracle_dense3c

    ref = oracle_dense3c(shells, aux, qvl=True)
    err = np.abs(got - ref)
    nz = np.abs(ref) > 1e-12
    _say(f"mixed blocks (200-atom recipe, 4 atoms): max abs {err.max():.2e}, max rel (|ref| > 1e-12) "
         f"{(err[nz] / np.abs(ref[nz])).max():.2e}, low-precision triplets {st.ncomputed_fp32} / {st.ncomputed}")
    assert st.ncomputed_fp32 > 0
    assert err.max() <= 2e-6 * opt.fp32_below


@pytest.mark.parametrize("key", ["w0", "w4r"])
def test_qvl_records_vs_oracle(key):
    """Contracted centres and extents (common/screening.f90:198-265),
    (min a + min b)^(1/4), Q_ab, and the aux extents and betas of the device
    path against the oracle's screener set-up."""
    if key == "w0":
        shells, aux, _ = synthetic.workload(0)
    else:
        shells, aux, _ = synthetic.workload4_reduced(10)
    b, x = _arrays(shells, aux)
    npk = len(shells.shells) * (len(shells.shells) + 1) // 2
    pr = np.zeros((npk, 8))
    xr = np.zeros((len(aux.shells), 4))
    N.check(N.lib().dfhip_qvl_records(ctypes.byref(b.struct), ctypes.byref(x.struct), pr.ctypes.data,
                                      xr.ctypes.data))
    bas, xb = O.Basis.from_shells(shells), O.Basis.from_shells(aux)
    po, xo = O.qvl_records(bas, xb, O.schwarz_bra(bas), O.schwarz_aux(xb))
    dc = np.abs(pr[:, :3] - po[:, :3]).max()
    de = np.abs(pr[:, 3] - po[:, 3]).max()
    _say(f"{key} QVL records: centre max abs diff {dc:.2e}, extent {de:.2e}")
    np.testing.assert_allclose(pr[:, :3], po[:, :3], rtol=0, atol=1e-13 * max(1.0, np.abs(po[:, :3]).max()))
    np.testing.assert_allclose(pr[:, 3], po[:, 3], rtol=1e-14)
    np.testing.assert_allclose(pr[:, 4], po[:, 4], rtol=1e-15)
    np.testing.assert_allclose(pr[:, 5], po[:, 5], rtol=1e-12)
    assert not pr[:, 6:].any()
    np.testing.assert_array_equal(xr[:, :2], xo[:, :2])
    np.testing.assert_allclose(xr[:, 2], xo[:, 2], rtol=1e-13)


def _device_densities(nstates, nbf, seed):
    """nstates(nstates+1)/2 symmetric (A + A^T)/2, A ~ U(0, 1), unit trace,
    drawn on the device (the 200-atom workload holds 102 GB of them)."""
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    ndens = nstates * (nstates + 1) // 2
    D = torch.empty((ndens, nbf, nbf), dtype=torch.float64, device=dev)
    for i in range(ndens):
        a = torch.rand((nbf, nbf), dtype=torch.float64, device=dev, generator=gen)
        a = 0.5 * (a + a.T)
        D[i] = a / torch.trace(a)
        del a
    torch.cuda.synchronize()
    return D


def _eightfold(J):
    perms = [(0, 1, 2, 3), (1, 0, 2, 3), (0, 1, 3, 2), (1, 0, 3, 2), (2, 3, 0, 1), (2, 3, 1, 0), (3, 2, 0, 1),
             (3, 2, 1, 0)]
    return all(np.array_equal(J, J.transpose(p)) for p in perms)


def test_workload4_full_properties():
    """The 200-atom workload (2,200 shells, 7,747,520,000 triplets, QVL 1e-10,
    20 states) on device densities: mixed within 1e-8 of FP64, the 8-fold
    symmetry of J, 64-bit counts, the class histogram summing to the
    screened count, both precisions screening the same triplets."""
    shells, aux, _ = synthetic.workload(4, with_densities=False)
    b, x = _arrays(shells, aux)
    nstates = 20
    _say(f"workload 4: nbf {shells.nbf} naux {aux.nbf}, drawing densities on the device ...")
    D = _device_densities(nstates, shells.nbf, synthetic.SEED_BASE + 4)
    res = {}
    try:
        for name, prec in (("fp64", N.FP64), ("mixed", N.MIXED)):
            J = np.zeros((nstates,) * 4, order="F")
            st = N.Stats()
            opt = N.default_options(precision=prec, screening=N.SCREEN_QVL, thresh=1e-10, dens_layout=N.DENS_C,
                                    dens_on_device=1)
            _say(f"workload 4 {name} ...")
            N.check(N.lib().dfhip_contract_coulomb_densities(
                ctypes.byref(b.struct), ctypes.byref(x.struct), D.shape[0], shells.nbf, D.data_ptr(),
                ctypes.byref(opt), None, nstates, J.ctypes.data, ctypes.byref(st)))
            res[name] = (J, st.as_dict())
            _say(f"workload 4 {name}: {st.ms_total:.0f} ms, computed {st.ncomputed} / {st.ntriplets}")
    finally:
        del D
        torch.cuda.empty_cache()
    (J64, s64), (Jm, sm) = res["fp64"], res["mixed"]
    assert s64["ntriplets"] == 7747520000
    for s in (s64, sm):
        _check_counts(s)
    assert sm["ncomputed"] == s64["ncomputed"]
    assert 0 < sm["ncomputed_fp32"] <= sm["ncomputed"] and s64["ncomputed_fp32"] == 0
    assert _eightfold(J64) and _eightfold(Jm)
    err = np.abs(Jm - J64).max()
    _say(f"workload 4: max |dJ| mixed vs FP64 {err:.2e} (|J|max {np.abs(J64).max():.3e}), low-precision "
         f"triplets {sm['ncomputed_fp32']} / {sm['ncomputed']}")
    assert err <= 1e-8


def test_batch_all_256_molecules(monkeypatch):
    """The batched workload in full: 256 molecules of 5-40 atoms (4 states)
    in one call, FP64 and mixed (with the FP32 windows of the generated
    kernels, DFHIP_GEN32=1: the default mixed run keeps these classes in
    FP64); per molecule the mixed J within 1e-8 of its FP64 J (which holds
    because those windows stay below 1e-3 under the default fp32_below of
    1e-2: at 1e-2 molecule 162 moves by 4.4e-8), 8-fold symmetry and
    consistent counts; the two smallest molecules against the oracle at
    1e-10."""
    raw, mols = [], []
    for i in range(256):
        sh, ax, dens = synthetic.batch_molecule(i)
        raw.append((sh, ax, dens))
        mols.append((N.BasisArrays(*sh.as_arrays()), N.BasisArrays(*ax.as_arrays()), dens))
    J64, st64 = _intor.contract_batch(mols, [4] * 256, options=N.default_options(precision=N.FP64))
    monkeypatch.setenv("DFHIP_GEN32", "1")
    Jm, stm = _intor.contract_batch(mols, [4] * 256, options=N.default_options(precision=N.MIXED))
    assert sum(s["ncomputed_fp32"] for s in stm) > 0
    worst = 0.0
    for k in range(256):
        _check_counts(st64[k])
        _check_counts(stm[k])
        assert stm[k]["ncomputed"] == st64[k]["ncomputed"]
        assert _eightfold(J64[k]) and _eightfold(Jm[k])
        d = np.abs(Jm[k] - J64[k]).max()
        worst = max(worst, d)
        assert d <= 1e-8, k
    _say(f"256 molecules: max |dJ| mixed vs FP64 {worst:.2e}, "
         f"{sum(s['ncomputed'] for s in st64)} screened triplets")
    for k in sorted(range(256), key=lambda k: raw[k][0].nbf)[:2]:
        Jref = O.contract_coulomb_densities_4d(*raw[k])
        err = np.abs(J64[k] - Jref).max()
        _say(f"molecule {k} (nbf {raw[k][0].nbf}): max |dJ| vs oracle {err:.2e}")
        assert err <= 1e-10
