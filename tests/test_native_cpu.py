"""Host-side checks of libdfhip.so that need no GPU: every symbol declared in
include/dfhip.h is exported, the Python shims validate their inputs like the
reference's typed memoryviews, and a call without a usable device reports an
error status instead of aborting (the reference aborts, common/normalization.f90:17-25)."""

import re
from pathlib import Path

import numpy as np
import pytest

from pysisyphus_addons_amd import _native as N
from pysisyphus_addons_amd import synthetic
from pysisyphus_addons_amd.shells import Shell, Shells

ROOT = Path(__file__).resolve().parents[1]


def _declared_symbols():
    text = (ROOT / "include" / "dfhip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(f_[a-z0-9_]+|dfhip_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_header_symbols():
    L = N.lib()
    declared = _declared_symbols()
    assert {"f_contract_coulomb_densities_2d", "f_contract_coulomb_densities_4d", "f_init",
            "f_int_schwarz_aux", "f_int_schwarz_bra"} <= declared
    for name in declared:
        assert hasattr(L, name), name
    assert set(N.EXPORTED) <= declared


def test_shell_arrays_roundtrip():
    shells, aux, _ = synthetic.molecule(3, 5.0, seed=1)
    bc, bs, bd = shells.as_arrays(fortran=True)
    assert bc.dtype == np.int32 and bs.dtype == np.int32 and bd.dtype == np.float64
    back = Shells.from_arrays(bc, bs, bd)
    assert len(back) == len(shells)
    for a, b in zip(back.shells, shells.shells):
        assert a.L == b.L
        np.testing.assert_array_equal(a.exps, b.exps)
        np.testing.assert_array_equal(a.coeffs, b.coeffs)
        np.testing.assert_array_equal(a.center, b.center)


def test_shim_rejects_wrong_dtype_and_layout():
    from pysisyphus_addons_amd.diabatization import _intor

    shells, aux, dens = synthetic.molecule(2, 4.0, nstates=1, seed=2)
    a = shells.as_arrays()
    x = aux.as_arrays()
    with pytest.raises(ValueError):
        _intor.contract_coulomb_densities_4d(a[0].astype(np.int64), a[1], a[2], *x, dens)
    with pytest.raises(ValueError):
        _intor.contract_coulomb_densities_4d(*a, *x, np.asfortranarray(dens))
    with pytest.raises(ValueError):
        _intor.contract_coulomb_densities_4d(*a, *x, dens.astype(np.float32))


def test_no_device_reports_status(monkeypatch):
    """In a process without a visible GPU the library must return an error
    status (no abort, no CPU fallback)."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from pysisyphus_addons_amd.diabatization import intor

    shells, aux, dens = synthetic.molecule(2, 4.0, nstates=1, seed=2)
    with pytest.raises(N.DfhipError) as e:
        intor.contract_coulomb_densities_4d(shells, aux, dens)
    assert e.value.status == 2
    # the batched form marshals its pointer arrays and reports the same status
    from pysisyphus_addons_amd.diabatization import _intor

    mols = [(N.BasisArrays(*shells.as_arrays()), N.BasisArrays(*aux.as_arrays()), dens)] * 3
    with pytest.raises(N.DfhipError) as e:
        _intor.contract_batch(mols, [1, 1, 1])
    assert e.value.status == 2


def test_workload_recipes():
    shells, aux, dens = synthetic.workload(0)
    assert len(shells) == 8 * 6 and shells.nbf == 112
    assert len(aux) == 8 * 10 and aux.nbf == 240
    assert dens.shape == (6, 112, 112)
    np.testing.assert_allclose(np.trace(dens, axis1=1, axis2=2), 1.0)
    np.testing.assert_allclose(dens, dens.transpose(0, 2, 1))
    ex = np.concatenate([s.exps for s in shells.shells])
    assert ex.min() >= 0.1 - 1e-12 and ex.max() <= 0.1 * 3 ** 7 + 1e-9


def test_invalid_basis_reports_status():
    """Basis errors are reported before any device work (the reference
    aborts the process, common/normalization.f90:17-25)."""
    from pysisyphus_addons_amd.diabatization import intor

    shells, aux, dens = synthetic.molecule(2, 4.0, nstates=1, seed=2)
    bad = Shells([Shell(L=6, center=[0, 0, 0], exps=[1.0], coeffs=[1.0])] + shells.shells[1:])
    with pytest.raises(N.DfhipError) as e:
        intor.contract_coulomb_densities_4d(bad, aux, dens)
    assert e.value.status == 4
    neg = Shells([Shell(L=0, center=[0, 0, 0], exps=[-1.0], coeffs=[1.0])] + shells.shells[1:])
    with pytest.raises(N.DfhipError) as e:
        intor.contract_coulomb_densities_4d(neg, aux, dens)
    assert e.value.status == 1
    from pysisyphus_addons_amd.diabatization import _intor

    a, x = shells.as_arrays(), aux.as_arrays()
    with pytest.raises(N.DfhipError) as e:
        _intor.contract_coulomb_densities_4d(*a, *x, np.ascontiguousarray(dens[:, :-1, :-1]))
    assert e.value.status == 1


def test_tuning_file_roundtrip(tmp_path):
    """The autotuner's table file (include/dfhip.h, dfhip_load_tuning) loads
    without a device, answers per class and window, and is rejected when
    malformed; loading None clears it."""
    from pysisyphus_addons_amd import tuning

    f = tmp_path / "tune.txt"
    f.write_text("# comment\n"
                 "int3c2e 2 1 3 0 2 0.5 0.25\n"
                 "int3c2e 2 2 2 1 1 0.1 0.2\n"
                 "contract 210 5 2.000 12.5\n"
                 "contract 36 103 1.000 1.5\n")
    try:
        tuning.load(str(f))
        t = tuning.table(210)
        assert t["int3c2e"] == {(2, 1, 3, 0): 2, (2, 2, 2, 1): 1}
        assert t["contract"] == (5, 2.0)
        assert tuning.table(36)["contract"] == (103, 1.0)
        assert tuning.table(20)["contract"] is None
        assert N.lib().dfhip_tuning_choice(1, 2, 0, 0) == -1  # lb > la
        bad = tmp_path / "bad.txt"
        bad.write_text("int3c2e 2 1 3 0 7\n")
        with pytest.raises(N.DfhipError, match="malformed"):
            tuning.load(str(bad))
        bad.write_text("contract 210 6 1.0\n")  # no tile of 6 densities per thread
        with pytest.raises(N.DfhipError, match="malformed"):
            tuning.load(str(bad))
        # a failed load keeps the previous table
        assert tuning.table()["int3c2e"] == {(2, 1, 3, 0): 2, (2, 2, 2, 1): 1}
    finally:
        tuning.load(None)
    assert tuning.table(210) == {"int3c2e": {}, "contract": None}


def _schwarz_bounds(shells, aux):
    b, x = N.BasisArrays(*shells.as_arrays()), N.BasisArrays(*aux.as_arrays())
    n = len(shells.shells)
    pb, xb = np.zeros(n * (n + 1) // 2), np.zeros(len(aux.shells))
    assert N.lib().dfhip_schwarz_bounds(b.struct, x.struct, pb.ctypes.data, xb.ctypes.data) == 0
    return pb, xb


@pytest.mark.parametrize("case", ["bena2", "w4r6", "spread"])
def test_schwarz_bounds_dominate_exact_norms(case, bena2):
    """The host bounds behind the contraction's pair drop (no GPU): every
    exact Q_ab and Q_C of the oracle lies below its bound; the bound is not
    vacuous (within 3x of Q for the tightest pair; 1 % above Q_C for a single
    aux primitive, within 2e3 for contracted aux shells);
    and far-apart pairs get bounds small enough to be dropped.
    bena2: contracted shells of a real molecule; w4r6: the 200-atom recipe
    (s-g, aux s-h) on 6 atoms; spread: 8 atoms of s-d shells over a 40 bohr
    cube, where most pairs are far apart."""
    import oracle as O

    if case == "bena2":
        shells = aux = Shells.from_arrays(bena2["bas_centers"], bena2["bas_spec"], bena2["bas_data"])
    elif case == "w4r6":
        shells, aux, _ = synthetic.workload4_reduced(6)
    else:
        shells, aux, _ = synthetic.molecule(8, 40.0, nstates=1, seed=7, with_densities=False)
    pb, xb = _schwarz_bounds(shells, aux)
    qb = O.schwarz_bra(O.Basis.from_shells(shells))
    qa = O.schwarz_aux(O.Basis.from_shells(aux))
    assert np.all(qb <= pb) and np.all(qa <= xb)
    live = qb > 1e-12
    assert np.min(pb[live] / qb[live]) < 3.0
    # a single aux primitive has the closed form (1 % above Q_C); a contracted
    # shell the primitive-wise bound, a fixed multiple per l (2.4 for s .. 1.4e3 for h)
    single = np.array([len(s.exps) == 1 for s in aux.shells])
    if single.any():
        r = xb[single] / qa[single]
        assert r.min() >= 1.0 and r.max() <= 1.0101, (r.min(), r.max())
    assert np.max(xb / qa) < 2e3
    if case == "spread":
        # pairs on different atoms 10+ bohr apart with all exponents >= 0.1
        assert np.mean(pb < 1e-14) > 0.5
