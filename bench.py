"""Benchmark: cells*PISO-steps/s of the PISO time step on one MI355X.

    python bench.py [--steps K] [--warmup W] [--config c3|c1] [--no-cpu-baseline]

Default workload (c3): the 2-block 256x256 periodic box (2 blocks of 256x128,
[0,2pi]^2, nu=1e-3, spectral k^-5/3 initial velocity, RMS 1, dt = 0.25 h/max|u0|),
FP32, BiCGStab advection / CG pressure at tol 1e-6, 2 correctors.  A step is one
PISO step over the whole domain; inputs are resident in HBM.

Prints one JSON line.  `roofline` is for the dominant kernel (the one with the
largest total time among the solver kernels, timed with HIP events on its launch
stream): algorithmic bytes per launch / average launch duration vs the 8 TB/s HBM
peak.  `cpu_baseline` is the FP32 C restatement of the same step (oracle/pict_oracle.c,
OpenMP) on this host's cores, over a bounded sample of steps of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def solver_bytes(kind, n, nnz, nrhs, s, iters):
    """Algorithmic HBM bytes of one solver launch (per-unit model of the PISO step's Krylov solvers):
    SpMV  = nnz*(s+4) + (n+1)*4 + 2*n*s
    CG iteration       = SpMV + 9*n*s            (x, r read/write; p, Ap read; r, p read/write)
    BiCGStab iteration = 2*SpMV + 17*n*s         (per right-hand side)
    plus the initial residual r0 = b - A x0 (SpMV + n*s for b) once per right-hand side."""
    spmv = nnz * (s + 4) + (n + 1) * 4 + 2 * n * s
    per_it = spmv + 9 * n * s if kind == "CG" else 2 * spmv + 17 * n * s
    return sum(max(it, 0) * per_it + spmv + n * s for it in iters)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c3", choices=["c3", "c1"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--copy-check", type=int, default=0, metavar="MIB",
                    help="after the timed region, run 3 plain device copies of MIB MiB (a known-size kernel that "
                         "calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the PISO step of one domain runs on a single MI355X "
              "(multi-GPU sharding is a design note, not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from pict_amd import PISOtorch as P
    from pict_amd import PISOtorch_simulation as S
    from pict_amd import configs

    dtype = torch.float32
    if args.config == "c3":
        dom, dt = configs.periodic_box_c3(256, dtype=dtype, device="cuda")
        workload = "periodic box 256x256, 2 blocks of 256x128, FP32, spectral k^-5/3 init, forward PISO step"
    else:
        dom = configs.lid_cavity(32, viscosity=0.01, dtype=dtype, device="cuda")
        dt = 0.01
        workload = "lid-driven cavity 32x32, Re=100, FP32, forward PISO step"
    N = dom.getTotalSize()
    sim = S.Simulation(dom, time_step=dt, corrector_steps=2, non_orthogonal=False, advection_tol=1e-6,
                       pressure_tol=1e-6, advect_passive_scalar=False)

    for _ in range(args.warmup):
        assert sim._PISO_split_step(1)
    torch.cuda.synchronize()
    P.enable_solve_log(True)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        assert sim._PISO_split_step(1)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = e0.elapsed_time(e1)
    log = P.solve_log
    P.enable_solve_log(False)

    # per-kernel roofline from the solver log
    agg = {}
    for kind, n, nnz, nrhs, s, infos, kms in log:
        a = agg.setdefault(kind, {"ms": 0.0, "bytes": 0, "launches": 0, "iters": 0})
        its = [i.usedIterations + 1 for i in infos]
        a["ms"] += kms
        a["bytes"] += solver_bytes(kind, n, nnz, nrhs, s, its)
        a["launches"] += 1
        a["iters"] += sum(its)
    dom_kind = max(agg, key=lambda k: agg[k]["ms"])
    d = agg[dom_kind]
    avg_ms = d["ms"] / d["launches"]
    bytes_per_launch = d["bytes"] / d["launches"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    ms_per_step = ms_total / args.steps
    value = N * args.steps / (ms_total * 1e-3)

    if args.copy_check > 0:
        src = torch.ones(args.copy_check * (1 << 18), dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(dom, dt, args)

    out = {
        "metric": "cells*PISO-steps/s (forward)",
        "value": value,
        "unit": "cells*steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 Krylov reductions)",
        "data": "synthetic (seeded)",
        "config": {"workload": workload, "cells": N, "config": args.config, "dt": dt,
                   "solver_iterations_per_step": {k: v["iters"] / args.steps for k, v in agg.items()},
                   "solver_ms_per_step": {k: v["ms"] / args.steps for k, v in agg.items()},
                   "wall_s": wall},
        "roofline": {"bound": "hbm", "kernel": "k_cg<float>" if dom_kind == "CG" else "k_bicgstab<float>",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_launch": bytes_per_launch, "avg_launch_ms": avg_ms},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def cpu_baseline(dom, dt, args):
    """FP32 C restatement of the same step (oracle/pict_oracle.c built with -fopenmp), on a sample of steps."""
    import torch
    from oracle import oracle as O
    cores = min(16, os.cpu_count() or 1)
    O.lib(32).oracle_set_threads(cores)
    twin = O.cpu_twin(dom, dtype=torch.float32)
    t0 = time.perf_counter()
    for _ in range(args.cpu_steps):
        O.piso_step(twin, dt, corrector_steps=2, advection_tol=1e-6, pressure_tol=1e-6)
    t = time.perf_counter() - t0
    return {"value": dom.getTotalSize() * args.cpu_steps / t, "unit": "cells*steps/s", "cores": cores,
            "kind": "port", "sample": "%d PISO steps of the same workload (FP32, OpenMP %d threads)"
            % (args.cpu_steps, cores)}


if __name__ == "__main__":
    main()
