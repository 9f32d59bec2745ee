"""The screening benchmark's records (intor/benchmark.f90:15-132, the C
symbol f_benchmark_int3c2e of intor/benchmark_wrapper.f90:12-38) and the
metric power of the eigen route (common/linalg.f90:36-58) against the FP64
oracle."""

import numpy as np
import pytest

import oracle as O
from pysisyphus_addons_amd import synthetic
from pysisyphus_addons_amd.intor import benchmark

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_estimate_records_vs_oracle():
    sh, ax, _ = synthetic.workload(0)
    got = benchmark.int3c2e_estimates(sh, ax)
    ref, nskip = O.int3c2e_estimates(sh, ax)
    assert got.shape == ref.shape
    _say(f"estimate records: {got.shape[0]}, max rel diff Schwarz "
         f"{np.max(np.abs(got[:, 0] - ref[:, 0]) / ref[:, 0]):.2e}, norms "
         f"{np.max(np.abs(got[:, 2] - ref[:, 2])):.2e} (max {ref[:, 2].max():.2e})")
    np.testing.assert_allclose(got[:, 0], ref[:, 0], rtol=1e-12)
    np.testing.assert_allclose(got[:, 1], ref[:, 1], rtol=1e-12)
    np.testing.assert_allclose(got[:, 2], ref[:, 2], rtol=1e-11, atol=1e-14 * ref[:, 2].max())
    np.testing.assert_allclose(got[:, 3], ref[:, 3], rtol=0, atol=1e-13)
    assert int(np.sum(got[:, 1] < O.SCHWARZ_THRESH)) == nskip


def test_benchmark_writes_int_estimates(tmp_path, monkeypatch, capfd):
    """f_benchmark_int3c2e writes the records to int_estimates.bin in the
    working directory (a stream of float64, intor/benchmark.f90:43,117) and
    logs the reference's lines."""
    sh, ax, _ = synthetic.workload(0)
    monkeypatch.chdir(tmp_path)
    benchmark.benchmark_int3c2e(sh, ax)
    out = capfd.readouterr().out
    recs = np.fromfile(tmp_path / "int_estimates.bin", dtype=np.float64).reshape(-1, 4)
    mem = benchmark.int3c2e_estimates(sh, ax)
    assert np.array_equal(recs, mem)
    assert "Skipped" in out and "Aux. basis" in out and "(P|Q)**-0.5" in out
    nsk = int(out.split("Skipped")[1].split()[0])
    assert nsk == int(np.sum(recs[:, 1] < O.SCHWARZ_THRESH))


@pytest.mark.parametrize("power", [-0.5, 1.0, -1.0])
def test_metric_powerh(power):
    """(P|Q)^power against the eigen decomposition of the oracle's metric
    (scipy): power 1 gives the metric back, -1 its inverse, -1/2 a matrix P
    with P M P = I; tolerances scale with the condition number."""
    from scipy.linalg import eigh

    _, ax, _ = synthetic.workload(0)
    M = O.metric(O.Basis.from_shells(ax))
    w, V = eigh(M)
    ref = (V * w ** power) @ V.T
    got = benchmark.metric_powerh(ax, power)
    kappa = w[-1] / w[0]
    err = np.abs(got - ref).max() / np.abs(ref).max()
    _say(f"(P|Q)^{power}: cond {kappa:.2e}, max rel diff {err:.2e}")
    assert err <= 1e-15 * kappa ** max(abs(power), 0.5) * 64
    if power == -0.5:
        I = got @ M @ got
        assert np.abs(I - np.eye(len(M))).max() <= 1e-13 * kappa
