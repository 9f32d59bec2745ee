"""Densities on grid points (grid/grid.f90:112-286) through f_eval_densities,
against the oracle's restatement of the blocked, screened evaluation:

* molecules with s-g shells, one and several densities, the reference's
  default block size (100) and ragged ones (37, 1, a block larger than the
  register group), thresholds that keep everything (0), the default (1e-8)
  and a coarse one (1e-4) that drops many pairs;
* a pair's kept/dropped decision is made on block means, so every test first
  checks on the oracle side that no pair sits within 1e-9 (relative) of the
  threshold, where rounding of the exponentials could flip a decision;
* the drop-in functions (eval_density, eval_densities), empty grids and the
  error statuses (Cartesian count, L = 5, block size 0).
The reference's own tests compare against pysisyphus' numba code on ORCA
wavefunctions that are not in the reference tree; parity here is anchored on
the oracle, which tests/test_oracle.py pins against analytic overlaps.
"""

import numpy as np
import pytest

import oracle as O
from pysisyphus_addons_amd import _native as N
from pysisyphus_addons_amd.grid import _grid, grid
from pysisyphus_addons_amd.shells import Shell, Shells

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _molecule(natoms, lmax, seed, side=4.0):
    rng = np.random.default_rng(seed)
    shells = []
    for ci in range(natoms):
        A = rng.uniform(0.0, side, 3)
        for L in range(lmax + 1):
            for _ in range(2 if L < 2 else 1):
                exps = np.sort(rng.uniform(0.15, 6.0, 3))[::-1]
                shells.append(Shell(L=L, center=A, exps=exps, coeffs=rng.uniform(0.2, 1.0, 3), center_ind=ci,
                                    atomic_num=6))
    return Shells(shells)


def _ncart(sh):
    return sum((s.L + 1) * (s.L + 2) // 2 for s in sh.shells)


def _densities(nc, ndens, seed):
    rng = np.random.default_rng(seed)
    D = rng.uniform(-1.0, 1.0, (ndens, nc, nc))
    return 0.5 * (D + D.transpose(0, 2, 1))


def _points(n, seed, lo=-2.0, hi=6.0):
    return np.random.default_rng(seed).uniform(lo, hi, (n, 3))


def _margin_ok(contribs, thresh):
    if thresh <= 0.0:
        return True
    worst = min(np.nanmin(np.abs(c / thresh - 1.0)) for c in contribs)
    return worst > 1e-9


CASES = [
    # natoms, lmax, npoints, ndens, blk_size, thresh
    (2, 2, 500, 1, 100, 1e-8),
    (3, 4, 1234, 2, 100, 1e-8),
    (3, 3, 777, 3, 37, 1e-4),
    (2, 4, 300, 1, 1, 1e-6),
    (3, 2, 1000, 2, 250, 1e-8),
    (2, 3, 401, 1, 100, 0.0),
]


@pytest.mark.parametrize("natoms,lmax,npoints,ndens,bs,thresh", CASES)
def test_grid_vs_oracle(natoms, lmax, npoints, ndens, bs, thresh):
    seed = 100 * natoms + 10 * lmax + ndens
    sh = _molecule(natoms, lmax, seed)
    nc = _ncart(sh)
    D = _densities(nc, ndens, seed + 1)
    pts = _points(npoints, seed + 2)
    arr = sh.as_arrays(fortran=True)
    ref, contribs = O.eval_densities(*arr, pts, D, blk_size=bs, thresh=thresh, return_masks=True)
    assert _margin_ok(contribs, thresh), "a pair sits at the threshold; pick another seed"
    kept = sum(int((c >= thresh).sum()) for c in contribs) if thresh > 0 else None
    total = sum(int(np.isfinite(c).sum()) for c in contribs)
    _say(f"grid {natoms} atoms lmax {lmax}: {npoints} points, {ndens} densities, blk {bs}, thresh {thresh:g}, "
         f"ncart {nc}; kept pairs {kept} / {total}")
    got = grid.eval_densities(sh, pts, D, blk_size=bs, thresh=thresh)
    assert got.shape == (ndens, npoints) and got.flags.f_contiguous
    err = np.abs(got - ref).max()
    _say(f"  max |d rho| {err:.2e} (max |rho| {np.abs(ref).max():.3e})")
    assert err <= 1e-12 * max(1.0, np.abs(ref).max())


def test_grid_single_density_and_layout():
    """eval_density (grid.py:4-10) and the raw _grid.eval_densities with the
    reference's defaults (blk_size 100, thresh 1e-8)."""
    sh = _molecule(2, 3, 7)
    nc = _ncart(sh)
    D = _densities(nc, 1, 8)
    pts = _points(333, 9)
    arr = sh.as_arrays(fortran=True)
    ref = O.eval_densities(*arr, pts, D)[0]
    rho = grid.eval_density(sh, pts, D[0])
    assert rho.shape == (333,)
    np.testing.assert_allclose(rho, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    raw = _grid.eval_densities(*arr, pts, D)
    np.testing.assert_array_equal(raw[0], rho)


def test_grid_many_blocks():
    """80,000 points (800 blocks) of a 6-atom s-g molecule, two densities:
    the block batching and the output indexing over a grid larger than the
    test cases above."""
    sh = _molecule(6, 4, 21, side=6.0)
    nc = _ncart(sh)
    D = _densities(nc, 2, 22)
    pts = _points(80000, 23, -3.0, 9.0)
    arr = sh.as_arrays(fortran=True)
    ref, contribs = O.eval_densities(*arr, pts, D, return_masks=True)
    assert _margin_ok(contribs, 1e-8)
    got = grid.eval_densities(sh, pts, D)
    err = np.abs(got - ref).max()
    _say(f"grid 80000 points, ncart {nc}: max |d rho| {err:.2e} (max |rho| {np.abs(ref).max():.3e})")
    assert err <= 1e-12 * max(1.0, np.abs(ref).max())


def test_grid_empty_and_errors():
    sh = _molecule(1, 2, 31)
    nc = _ncart(sh)
    D = _densities(nc, 2, 32)
    out = grid.eval_densities(sh, np.zeros((0, 3)), D)
    assert out.shape == (2, 0)
    pts = _points(10, 33)
    with pytest.raises(N.DfhipError, match="Cartesian"):
        grid.eval_densities(sh, pts, _densities(nc - 1, 1, 34))
    with pytest.raises(N.DfhipError, match="blk_size"):
        grid.eval_densities(sh, pts, D, blk_size=0)
    h = Shells([Shell(L=5, center=[0.0, 0.0, 0.0], exps=[1.0], coeffs=[1.0])])
    with pytest.raises(N.DfhipError) as e:
        grid.eval_densities(h, pts, _densities(21, 1, 35))
    assert e.value.status == 4  # DFHIP_ERR_UNSUPPORTED
