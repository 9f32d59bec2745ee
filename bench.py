"""Benchmark of the density-fitting Coulomb path on one MI355X.

A step is one complete Coulomb-tensor evaluation (Schwarz norms, screener
records, metric, Cholesky, screened 3c2e integrals contracted with the
densities, B = L^-1 d, the 8-fold J assembly).  The default workload is the
200-atom system (34 bohr cube, orbital 4s3p2d1f1g of 3-primitive
contractions, aux 5s4p3d2f1g1h uncontracted, QVL screening at 1e-10, 20
states = 210 densities drawn on the device), with basis and densities
resident in HBM when the timed region starts, the densities in the layout
the reference's Cython passes (Fortran order of (ndens, nbf, nbf), the
density index fastest; `--dens-layout c` for row-major).  The same step runs on the
plain FP64 path (the drop-in) and on the mixed-precision path.

`--workload 1` and `--workload 2` run the 60-atom systems (orbital s-d / aux
s-f and orbital s-f / aux s-g, 8 states); `--batch` runs 256 independent
molecules in one batched call.

Prints one JSON line: value = screened shell-triplets per second of the whole
mixed-precision step (ms_per_step its wall time), `fp64` = the same for the
FP64 path, speedup_vs_fp64 and the largest deviation of the mixed J from the
FP64 J; roofline of the 3c2e kernels of the value's path (algorithmic flops
from pysisyphus_addons_amd.flops over their HIP-event time on the library
stream); cpu_baseline = the FP64 oracle's integral-direct loop on a sample of
the same workload on this host's cores.
"""

import argparse
import ctypes
import json
import os
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "pysisyphus-addons_amd"))

METRIC = "screened 3c2e shell-triplets/s + J-tensor time: mixed vs FP64 GPU at ≤1e-8 Eh"
WORKLOADS = {
    1: "60 atoms in a 22 bohr cube, orbital 3s2p1d (3-prim), aux 4s3p2d1f, 8 states, Schwarz 1e-10",
    2: "60 atoms in a 22 bohr cube, orbital 3s2p1d1f (3-prim), aux 4s3p2d1f1g, 8 states, Schwarz 1e-10",
    4: "200 atoms in a 34 bohr cube, orbital 4s3p2d1f1g (3-prim, exponents 0.05-2000), aux 5s4p3d2f1g1h "
       "(0.03-500), 20 states, QVL 1e-10",
}
FP64_PEAK_TFLOPS = 78.6
# rocprofv3 FETCH_SIZE / WRITE_SIZE summaries of the 3c2e kernels, per workload and path
TRAFFIC = {
    (1, "fp64"): ROOT / "profiles" / "traffic_60atom_fp64.md",
    (4, "fp64"): ROOT / "profiles" / "traffic_200atom_fp64.md",
    (4, "mixed"): ROOT / "profiles" / "traffic_200atom_mixed.md",
}


def measured_traffic(workload, path):
    """HBM bytes of the 3c2e kernels per step from the committed rocprofv3
    summary of this workload and path (None without one)."""
    f = TRAFFIC.get((workload, path))
    try:
        for line in f.read_text().splitlines():
            if line.startswith("int3c2e_hbm_bytes_per_step:"):
                return int(line.split(":")[1]), str(f.relative_to(ROOT))
    except (OSError, AttributeError):
        pass
    return None, None


def host_cpu():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        info = dict(line.split(":", 1) for line in out.splitlines() if ":" in line)
        return {k: info.get(k, "").strip() for k in ("Model name", "Socket(s)", "Core(s) per socket",
                                                     "Thread(s) per core", "CPU(s)")}
    except Exception:
        return {}


def cpu_baseline(shells, aux, wid, screening, ncomputed_full):
    """The oracle's integral-direct loop (diabatization/intor.f90:108-183
    restated, FP64, OpenMP over aux shells) on a bounded sample of the
    workload.  Workloads 1 and 2: every aux shell.  Workload 4: every triplet
    among the K atoms nearest the cube centre (their orbital shell pairs x
    their aux shells), screened with the workload's own estimates, which
    depend only on the pair and the aux shell (Schwarz norms from the
    library's f_int_schwarz_bra / f_int_schwarz_aux)."""
    import numpy as np

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(os.cpu_count() or 1, 16)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O

    from pysisyphus_addons_amd import synthetic
    from pysisyphus_addons_amd.intor import intor as sintor
    from pysisyphus_addons_amd.shells import Shells

    rng = np.random.Generator(np.random.PCG64(7))
    if wid == 4:
        K = 16
        cents = np.array([s.center for s in shells.shells])
        atoms = sorted({s.center_ind for s in shells.shells})
        pos = {s.center_ind: s.center for s in shells.shells}
        mid = cents.mean(axis=0)
        near = sorted(atoms, key=lambda a: np.linalg.norm(pos[a] - mid))[:K]
        sub = Shells([s for s in shells.shells if s.center_ind in near])
        subx = Shells([s for s in aux.shells if s.center_ind in near])
        desc = (f"all triplets among the {K} atoms nearest the cube centre ({len(sub.shells)} orbital shells, "
                f"{len(subx.shells)} aux shells, QVL 1e-10)")
    else:
        sub, subx = shells, aux
        desc = "all aux shells (Schwarz 1e-10)"
    nst = synthetic.WORKLOADS[wid]["nstates"]
    dens = synthetic.random_densities(rng, nst, sub.nbf)
    bas, x = O.Basis.from_shells(sub), O.Basis.from_shells(subx)
    qb = sintor.int_schwarz_bra(sub) if wid == 4 else O.schwarz_bra(bas)
    qa = sintor.int_schwarz_aux(subx) if wid == 4 else O.schwarz_aux(x)
    qvl = O.qvl_records(bas, x, qb, qa) if screening == "qvl" else None
    t0 = time.perf_counter()
    _, nskip, ncomp = O.df_contract(bas, x, dens, qbra=qb, qaux=qa, thresh=1e-10 if wid == 4 else O.SCHWARZ_THRESH,
                                    qvl=qvl)
    dt = time.perf_counter() - t0
    out = {
        "value": ncomp / dt,
        "unit": "screened shell-triplets/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{desc}: {ncomp} screened triplets in {dt:.1f} s, 3c2e + density contraction loop",
        "host": host_cpu(),
    }
    if wid == 4:
        out["extrapolated"] = True
        out["extrapolated_contraction_s"] = ncomputed_full * dt / max(ncomp, 1)
    return out


def mfma_calibration():
    """One grid-density call (4 atoms with s-d shells, 56 Cartesian
    functions, 10,000 points in blocks of 100, one density, thresh 0) and the
    exact number of v_mfma_f64_16x16x4f64 its density kernel issues: with
    thresh 0 no 4 x 16 slice of the density is screened out, so per block and
    64-point group it runs ceil((ncart - 16 t) / 4) K steps for every
    16-column tile t, each on every 16-point row tile of the group."""
    import numpy as np

    from pysisyphus_addons_amd.grid import grid
    from pysisyphus_addons_amd.shells import Shell, Shells

    rng = np.random.default_rng(5)
    shs = []
    for ci in range(4):
        A = rng.uniform(0.0, 6.0, 3)
        for L in (0, 0, 1, 1, 2):
            shs.append(Shell(L=L, center=A, exps=np.sort(rng.uniform(0.15, 6.0, 3))[::-1],
                             coeffs=rng.uniform(0.2, 1.0, 3), center_ind=ci))
    sh = Shells(shs)
    nc = sum((s.L + 1) * (s.L + 2) // 2 for s in sh.shells)
    D = rng.uniform(-1.0, 1.0, (1, nc, nc))
    D = 0.5 * (D + D.transpose(0, 2, 1))
    pts = rng.uniform(-2.0, 8.0, (10000, 3))
    bs = 100
    grid.eval_densities(sh, pts, D, blk_size=bs, thresh=0.0)
    n = 0
    for b0 in range(0, len(pts), bs):
        cur = min(bs, len(pts) - b0)
        for g0 in range(0, cur, 64):
            nit = (min(64, cur - g0) + 15) // 16
            n += nit * sum((nc - 16 * t + 3) // 4 for t in range((nc + 15) // 16))
    print(json.dumps({"kernel": "grid_density_kernel", "ncart": nc, "mfma_f64_16x16x4": n, "flops": 2048 * n}),
          flush=True)


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
    dtm, Jm, ncm = res["mixed"]
    out = {
        "metric": METRIC,
        "value": ncm / dtm,
        "unit": "screened shell-triplets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dtm * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/f64 mixed",
        "data": "synthetic",
        "config": {"workload": "256 molecules of 5-40 atoms (cube side 4 N^(1/3) bohr), orbital 3s2p1d, "
                               "aux 4s3p2d1f, 4 states each, one batched call",
                   "screened_triplets": ncm},
        "fp64": {"value": ncomp / dt, "ms_per_step": dt * 1e3},
        "speedup_vs_fp64": dt / dtm,
        "max_abs_dJ_vs_fp64": float(max(np.abs(a - b).max() for a, b in zip(Jm, J64))),
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="only the FP64 path (value is then FP64)")
    ap.add_argument("--no-fp64", action="store_true", help="only the mixed path (no speed-up, no dJ)")
    ap.add_argument("--fp32-below", type=float, default=None, help="mixed path: low precision below this "
                                                                     "Schwarz product")
    ap.add_argument("--batch", action="store_true",
                    help="the batched form instead: 256 molecules of 5-40 atoms, 4 states, one call")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only a 2 GiB device-to-device copy: the reference run that scales the FETCH_SIZE and "
                         "WRITE_SIZE counters of the roofline's traffic figure")
    ap.add_argument("--mfma-calibration", action="store_true",
                    help="only one f_eval_densities call whose density kernel issues a known number of FP64 "
                         "MFMA instructions (printed): the reference run for the unit of "
                         "SQ_INSTS_VALU_MFMA_MOPS_F64")
    ap.add_argument("--dens-layout", choices=("c", "fortran"), default="fortran",
                    help="layout of the device-resident densities: c = (ndens, nbf, nbf) row-major, fortran = "
                         "(ndens, nbf, nbf) column-major (what the reference's Cython passes)")
    ap.add_argument("--workload", type=int, default=4, choices=(1, 2, 4),
                    help="4: 200 atoms s-g / aux s-h, 20 states, QVL 1e-10, densities generated on the device "
                         "(default); 1: 60 atoms s-d / aux s-f, 8 states; 2: 60 atoms s-f / aux s-g, 8 states")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the path shards over auxiliary shells across GPUs by design, but only the one-GPU share is "
              "built and measured here)", file=sys.stderr)
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
    if args.mfma_calibration:
        return mfma_calibration()
    if args.batch:
        return bench_batch(args)

    dev = torch.device("cuda:0")
    fortran = args.dens_layout == "fortran"
    base_opt = dict(dens_layout=N.DENS_FORTRAN if fortran else N.DENS_C, dens_on_device=1)
    wid = args.workload
    nstates = synthetic.WORKLOADS[wid]["nstates"]
    if wid == 4:
        # 210 densities of 7800^2 (102 GB): drawn on the device, (A + A^T)/2 with
        # A ~ U(0, 1), unit trace
        shells, aux, _ = synthetic.workload(4, with_densities=False)
        nbf, ndens = shells.nbf, nstates * (nstates + 1) // 2
        gen = torch.Generator(device=dev)
        gen.manual_seed(synthetic.SEED_BASE + 4)
        # C layout: [I][i][j]; Fortran layout of (ndens, nbf, nbf): the density
        # index fastest, element (I, i, j) at I + ndens (i + nbf j), which is
        # [j][i][I] of a C-contiguous (nbf, nbf, ndens) tensor
        dens_d = torch.empty((nbf, nbf, ndens) if fortran else (ndens, nbf, nbf), dtype=torch.float64, device=dev)
        for i in range(ndens):
            a = torch.rand((nbf, nbf), dtype=torch.float64, device=dev, generator=gen)
            a = 0.5 * (a + a.T)
            if fortran:
                dens_d[:, :, i] = a / torch.trace(a)  # symmetric: [j][i] = [i][j]
            else:
                dens_d[i] = a / torch.trace(a)
            del a
        base_opt.update(screening=N.SCREEN_QVL, thresh=1e-10)
        screening = "qvl"
    else:
        shells, aux, dens = synthetic.workload(wid)
        ndens, nbf = dens.shape[0], dens.shape[1]
        dens_d = torch.from_numpy(np.ascontiguousarray(dens.transpose(2, 1, 0)) if fortran else dens).to(dev)
        screening = "schwarz"
    b = N.BasisArrays(*shells.as_arrays())
    x = N.BasisArrays(*aux.as_arrays())
    torch.cuda.synchronize()
    J = np.zeros((nstates,) * 4, order="F")
    st = N.Stats()

    def run(opt, label):
        def step():
            s = N.lib().dfhip_contract_coulomb_densities(
                ctypes.byref(b.struct), ctypes.byref(x.struct), ndens, nbf, dens_d.data_ptr(),
                ctypes.byref(opt), None, nstates, J.ctypes.data, ctypes.byref(st))
            N.check(s)
            # a raw copy of the stats: the conversion to a dict waits until
            # the timed loop is over
            return N.Stats.from_buffer_copy(st)

        for i in range(args.warmup):
            step()
            print(f"{label}: warmup step {i + 1} done", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        recs = []
        t0 = time.perf_counter()
        for i in range(args.steps):
            recs.append(step())
            print(f"{label}: step {i + 1} done", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        recs = [r.as_dict() for r in recs]
        ms = dt / args.steps * 1e3
        r0 = recs[-1]
        avg = lambda k: float(np.mean([r[k] for r in recs]))
        ms_int = avg("ms_int3c2e")
        alg_flops = flops.total_flops(r0["class_prim_triplets"])
        res = {
            "value": r0["ncomputed"] / (ms * 1e-3),
            "ms_per_step": ms,
            "j_tensor_ms": ms,
            "int3c2e_screened_triplets_per_s": r0["ncomputed"] / ((ms_int + avg("ms_contract")) * 1e-3),
            "phases_ms": {k[3:]: avg(k) for k in r0 if k.startswith("ms_")},
            "flops_per_launch": alg_flops,
            "ms_per_launch": ms_int,
            "achieved": alg_flops / (ms_int * 1e-3) / 1e12,
        }
        print(f"{label}: {ms:.2f} ms per step", file=sys.stderr, flush=True)
        return res, r0, J.copy()

    paths = {}
    if not args.no_fp64:
        paths["fp64"] = run(N.default_options(**base_opt), "fp64")
    if not args.no_mixed:
        mopt = N.default_options(precision=N.MIXED, **base_opt)
        if args.fp32_below is not None:
            mopt.fp32_below = args.fp32_below
        res, r0, Jm = run(mopt, "mixed")
        res["fp32_below"] = mopt.fp32_below
        res["low_precision_triplets"] = r0["ncomputed_fp32"]
        paths["mixed"] = (res, r0, Jm)
    head = "mixed" if "mixed" in paths else "fp64"
    res, r0, Jh = paths[head]
    traffic, tsrc = measured_traffic(wid, head)
    out = {
        "metric": METRIC,
        "value": res["value"],
        "unit": "screened shell-triplets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/f64 mixed" if head == "mixed" else "f64",
        "data": "synthetic",
        "config": {
            "workload": WORKLOADS[wid],
            "path": head,
            "density_layout": args.dens_layout,
            "nbf": nbf,
            "naux": aux.nbf,
            "ndens": ndens,
            "screened_triplets": r0["ncomputed"],
            "triplets": r0["ntriplets"],
            "j_tensor_ms": res["ms_per_step"],
            "int3c2e_screened_triplets_per_s": res["int3c2e_screened_triplets_per_s"],
            "phases_ms": res["phases_ms"],
            "class_triplets": r0["class_triplets"],
        },
        "roofline": {
            "bound": "mfma",
            "achieved": res["achieved"],
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": res["achieved"] / FP64_PEAK_TFLOPS,
            "traffic": traffic,
            "traffic_source": tsrc,
            "kernel": "3c2e class kernels (all classes of one step, int3c2e_mm_kernel the largest family)",
            "flops_per_launch": res["flops_per_launch"],
            "ms_per_launch": res["ms_per_launch"],
        },
    }
    if head == "mixed":
        out["mixed"] = {k: v for k, v in res.items() if k in ("fp32_below", "low_precision_triplets")}
    if "fp64" in paths and head == "mixed":
        f64 = paths["fp64"][0]
        out["fp64"] = {k: f64[k] for k in ("value", "ms_per_step", "int3c2e_screened_triplets_per_s",
                                             "phases_ms", "achieved", "ms_per_launch")}
        out["speedup_vs_fp64"] = f64["ms_per_step"] / res["ms_per_step"]
        out["max_abs_dJ_vs_fp64"] = float(np.abs(Jh - paths["fp64"][2]).max())
    if not args.no_cpu_baseline:
        del dens_d
        torch.cuda.empty_cache()
        out["cpu_baseline"] = cpu_baseline(shells, aux, wid, screening, r0["ncomputed"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
