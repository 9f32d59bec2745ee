"""Benchmark of the density-fitting Coulomb path on one MI355X.

A step is one complete Coulomb-tensor evaluation (Schwarz norms, metric,
Cholesky, screened 3c2e integrals contracted with the densities, B = L^-1 d,
the 8-fold J assembly) for the 60-atom workload (22 bohr cube, orbital
3s2p1d, aux 4s3p2d1f, 8 states = 36 densities, FP64), with basis and
densities resident in HBM when the timed region starts.

`--workload 2` and `--workload 4` run the same step on the f/g-orbital
60-atom system and on the 200-atom g/h system (QVL screening at 1e-10, 20
states, densities drawn on the device); `--batch` runs 256 independent
molecules in one batched call.

Prints one JSON line: value = screened shell-triplets per second of the
whole step; roofline of the 3c2e kernels (algorithmic FP64 flops from
pysisyphus_addons_amd.flops over their HIP-event time on the library stream);
mixed = the same step with the mixed-precision path (FP32 integral arithmetic
for triplets with Schwarz product below 1e-3), its throughput and the largest
deviation of its J from the FP64 path's; cpu_baseline = the FP64 oracle's
integral-direct loop on a sample of the same workload on this host's cores.
"""

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "pysisyphus-addons_amd"))

METRIC = "screened 3c2e shell-triplets/s + J-tensor time: mixed vs FP64 GPU at ≤1e-8 Eh"
WORKLOADS = {
    1: "60 atoms in a 22 bohr cube, orbital 3s2p1d (3-prim), aux 4s3p2d1f, 8 states, FP64",
    2: "60 atoms in a 22 bohr cube, orbital 3s2p1d1f (3-prim), aux 4s3p2d1f1g, 8 states, FP64",
    4: "200 atoms in a 34 bohr cube, orbital 4s3p2d1f1g (3-prim), aux 5s4p3d2f1g1h, 20 states, QVL 1e-10, FP64",
}
FP64_PEAK_TFLOPS = 78.6
TRAFFIC_SUMMARY = ROOT / "profiles" / "traffic_60atom_fp64.md"


def measured_traffic():
    """HBM bytes of the 3c2e kernels per step, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload (None without it)."""
    try:
        for line in TRAFFIC_SUMMARY.read_text().splitlines():
            if line.startswith("int3c2e_hbm_bytes_per_step:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def cpu_baseline(shells, aux, dens, step):
    """Oracle integral-direct loop over every `step`-th aux shell."""
    threads = min(os.cpu_count() or 1, 16)
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O

    bas, x = O.Basis.from_shells(shells), O.Basis.from_shells(aux)
    qb, qa = O.schwarz_bra(bas), O.schwarz_aux(x)
    t0 = time.perf_counter()
    _, nskip, ncomp = O.df_contract(bas, x, dens, qbra=qb, qaux=qa, c_begin=0, c_end=x.nshells,
                                    c_step=step)
    dt = time.perf_counter() - t0
    return {
        "value": ncomp / dt,
        "unit": "screened shell-triplets/s",
        "cores": int(os.environ.get("OMP_NUM_THREADS", threads)),
        "kind": "port",
        "sample": (("all aux shells" if step == 1 else f"every {step}th aux shell")
                   + f" ({ncomp} screened triplets, {dt:.1f} s) of the same "
                   "workload, 3c2e + density contraction loop"),
    }


def bench_batch(args):
    """Batched form: 256 independent molecules (5-40 atoms, cube side
    4 N^(1/3), orbital 3s2p1d, aux 4s3p2d1f, 4 states) per step, FP64 and
    mixed, through dfhip_contract_coulomb_densities_batch."""
    import numpy as np

    from pysisyphus_addons_amd import _native as N
    from pysisyphus_addons_amd import synthetic
    from pysisyphus_addons_amd.diabatization import _intor

    mols = []
    for i in range(256):
        sh, ax, dens = synthetic.batch_molecule(i)
        mols.append((N.BasisArrays(*sh.as_arrays()), N.BasisArrays(*ax.as_arrays()), dens))
    res = {}
    for name, prec in (("fp64", N.FP64), ("mixed", N.MIXED)):
        opt = N.default_options(precision=prec)
        for _ in range(args.warmup):
            _intor.contract_batch(mols, [4] * 256, options=opt)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            Js, sts = _intor.contract_batch(mols, [4] * 256, options=opt)
        dt = (time.perf_counter() - t0) / args.steps
        res[name] = (dt, Js, sum(st["ncomputed"] for st in sts))
    dt, J64, ncomp = res["fp64"]
    out = {
        "metric": METRIC,
        "value": ncomp / dt,
        "unit": "screened shell-triplets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "256 molecules of 5-40 atoms (cube side 4 N^(1/3) bohr), orbital 3s2p1d, "
                               "aux 4s3p2d1f, 4 states each, one batched call on 4 lanes",
                   "screened_triplets": ncomp},
        "mixed": {"value": res["mixed"][2] / res["mixed"][0], "ms_per_step": res["mixed"][0] * 1e3,
                  "max_abs_dJ_vs_fp64": float(max(np.abs(a - b).max() for a, b in zip(res["mixed"][1], J64)))},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-step", type=int, default=1, help="aux-shell stride of the CPU sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision measurement")
    ap.add_argument("--fp32-below", type=float, default=None, help="mixed path: FP32 below this Schwarz product")
    ap.add_argument("--batch", action="store_true",
                    help="the batched form instead: 256 molecules of 5-40 atoms, 4 states, one call")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only a 2 GiB device-to-device copy: the reference run that scales the FETCH_SIZE and "
                         "WRITE_SIZE counters of the roofline's traffic figure")
    ap.add_argument("--workload", type=int, default=1, choices=(1, 2, 4),
                    help="1: 60 atoms s-d / aux s-f, 8 states (default); 2: 60 atoms s-f / aux s-g, 8 states; "
                         "4: 200 atoms s-g / aux s-h, 20 states, QVL 1e-10, densities generated on the device")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the path shards over auxiliary shells, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from pysisyphus_addons_amd import _native as N
    from pysisyphus_addons_amd import flops, synthetic

    if args.copy_calibration:
        x = torch.ones(2 ** 28, dtype=torch.float64, device="cuda:0")
        y = torch.empty_like(x).copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copied_bytes": x.numel() * 8, "check": float(y[-1])}), flush=True)
        return None
    if args.batch:
        return bench_batch(args)

    dev = torch.device("cuda:0")
    base_opt = dict(dens_layout=N.DENS_C, dens_on_device=1)
    if args.workload == 4:
        # 210 densities of 7800^2 (102 GB): drawn on the device, (A + A^T)/2 with
        # A ~ U(0, 1), unit trace
        shells, aux, _ = synthetic.workload(4, with_densities=False)
        nstates = synthetic.WORKLOADS[4]["nstates"]
        nbf, ndens = shells.nbf, nstates * (nstates + 1) // 2
        gen = torch.Generator(device=dev)
        gen.manual_seed(synthetic.SEED_BASE + 4)
        dens_d = torch.empty((ndens, nbf, nbf), dtype=torch.float64, device=dev)
        for i in range(ndens):
            a = torch.rand((nbf, nbf), dtype=torch.float64, device=dev, generator=gen)
            a = 0.5 * (a + a.T)
            dens_d[i] = a / torch.trace(a)
            del a
        dens = None
        base_opt.update(screening=N.SCREEN_QVL, thresh=1e-10)
    else:
        shells, aux, dens = synthetic.workload(args.workload)
        ndens, nbf = dens.shape[0], dens.shape[1]
        nstates = synthetic.WORKLOADS[args.workload]["nstates"]
        dens_d = torch.from_numpy(dens).to(dev)
    b = N.BasisArrays(*shells.as_arrays())
    x = N.BasisArrays(*aux.as_arrays())
    torch.cuda.synchronize()
    J = np.zeros((nstates,) * 4, order="F")
    st = N.Stats()

    def run(opt):
        def step():
            s = N.lib().dfhip_contract_coulomb_densities(
                ctypes.byref(b.struct), ctypes.byref(x.struct), ndens, nbf, dens_d.data_ptr(),
                ctypes.byref(opt), None, nstates, J.ctypes.data, ctypes.byref(st))
            N.check(s)
            # a raw copy of the stats: the conversion to a dict waits until
            # the timed loop is over
            return N.Stats.from_buffer_copy(st)

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        recs = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            recs.append(step())
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        return dt / args.steps * 1e3, [r.as_dict() for r in recs], J.copy()

    ms_per_step, recs, J64 = run(N.default_options(**base_opt))
    print(f"fp64: {ms_per_step:.2f} ms per step", file=sys.stderr, flush=True)
    mixed = None
    if not args.no_mixed:
        mopt = N.default_options(precision=N.MIXED, **base_opt)
        if args.fp32_below is not None:
            mopt.fp32_below = args.fp32_below
        ms_mixed, mrecs, Jm = run(mopt)
        print(f"mixed: {ms_mixed:.2f} ms per step", file=sys.stderr, flush=True)
        m0 = mrecs[-1]
        mixed = {
            "value": m0["ncomputed"] / (ms_mixed * 1e-3),
            "ms_per_step": ms_mixed,
            "speedup_vs_fp64": ms_per_step / ms_mixed,
            "fp32_below": mopt.fp32_below,
            "fp32_triplets": m0["ncomputed_fp32"],
            "max_abs_dJ_vs_fp64": float(np.abs(Jm - J64).max()),
            "phases_ms": {k[3:]: float(np.mean([r[k] for r in mrecs])) for k in m0 if k.startswith("ms_")},
        }
    r0 = recs[-1]
    ncomp = r0["ncomputed"]
    avg = lambda k: float(np.mean([r[k] for r in recs]))
    ms_int = avg("ms_int3c2e")
    alg_flops = flops.total_flops(r0["class_prim_triplets"])
    achieved = alg_flops / (ms_int * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": ncomp / (ms_per_step * 1e-3),
        "unit": "screened shell-triplets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": WORKLOADS[args.workload],
            "nbf": nbf,
            "naux": aux.nbf,
            "ndens": ndens,
            "screened_triplets": ncomp,
            "triplets": r0["ntriplets"],
            "precision": "fp64",
            "j_tensor_ms": ms_per_step,
            "int3c2e_screened_triplets_per_s": ncomp / ((ms_int + avg("ms_contract")) * 1e-3),
            "phases_ms": {k[3:]: avg(k) for k in r0 if k.startswith("ms_")},
        },
        "roofline": {
            "bound": "mfma",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": measured_traffic(),
            "traffic_source": "profiles/traffic_60atom_fp64.md (rocprofv3 FETCH_SIZE, WRITE_SIZE, copy-calibrated)",
            "kernel": "int3c2e class kernels (all classes of one step)",
            "flops_per_launch": alg_flops,
            "ms_per_launch": ms_int,
        },
    }
    if mixed is not None:
        out["mixed"] = mixed
    if not args.no_cpu_baseline and args.workload == 1:
        out["cpu_baseline"] = cpu_baseline(shells, aux, dens, args.cpu_step)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
