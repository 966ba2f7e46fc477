"""Benchmark: cells*PISO-steps/s of the PISO time step on one MI355X.

    python bench.py [--mode fwdbwd|fwd] [--steps K] [--warmup W] [--config c3|c1|c4] [--no-cpu-baseline]

Default workload (differentiation configuration): the 2-block 256x256 periodic box (2 blocks of 256x128,
[0,2pi]^2, nu=1e-3, spectral k^-5/3 initial velocity with RMS 1, dt = 0.25 h/max|u0|), FP32, BiCGStab advection
and CG pressure at tol 1e-6, 2 correctors; a learnable per-cell forcing (block velocity sources, zero-initialised);
a K-step rollout (default 16) with per-step checkpointing, MSE loss to a seeded target state, and backward.
A bench step is one PISO step forward + its backward (recomputation included); the timed region is one whole
rollout plus backward, inputs resident in HBM.  `--mode fwd` times K forward steps only.

`--config c4` (batching configuration): 64 independent lid cavities (sides uniform over 32..256, viscosities
log-uniform over 1e-3..1e-1), K = 100 forward steps each, stepped as one batch (one launch per stage and per
solve, graph replay, autotuned launch shapes) and, for comparison, as the plain loop of one native step per
cavity; `value` is the batch's, `config.batched_vs_plain` the ratio.  The default run appends the same measurement
(without its CPU sample) to its line under `batched` (`--no-batched` skips it).

`config.error_vs_fp64` states the FP32 error against FP64 steps of the same workload: after the whole forward
trajectory from the same initial state, and per step, from FP64 states along the trajectory, against the FP64
step whose solves run the FP32 step's iteration counts (at tolerance 1e-6 a solve near the tolerance may stop one
iteration earlier or later in either precision, which alone moves the state by up to about 1e-3).

Prints one JSON line.  `roofline` is for the dominant kernel (the solver kernel with the largest total time in the
timed region, timed with HIP events on its launch stream): algorithmic bytes per launch / average launch duration
vs the 8 TB/s HBM peak.  `cpu_baseline` is the FP32 C restatement (oracle/pict_oracle.c, OpenMP) of the same
step -- forward, or forward plus the reverse pass -- on this host's cores, over a bounded sample of steps.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
_T0 = time.perf_counter()
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def _note(msg):
    """A progress line on stderr (the JSON line stays alone on stdout)."""
    print("[bench %.0fs] %s" % (time.perf_counter() - _T0, msg), file=sys.stderr, flush=True)


def solver_bytes(kind, n, nnz, nrhs, s, iters):
    """Algorithmic HBM bytes of one solver launch (per-unit model of the PISO step's Krylov solvers):
    SpMV  = nnz*(s+4) + (n+1)*4 + 2*n*s
    CG iteration       = SpMV + 9*n*s            (x, r read/write; p, Ap read; r, p read/write)
    BiCGStab iteration = 2*SpMV + 17*n*s         (per right-hand side)
    plus the initial residual r0 = b - A x0 (SpMV + n*s for b) once per right-hand side.  Adjoint (transposed)
    solves are counted like forward ones."""
    spmv = nnz * (s + 4) + (n + 1) * 4 + 2 * n * s
    per_it = spmv + 9 * n * s if kind == "CG" else 2 * spmv + 17 * n * s
    return sum(max(it, 0) * per_it + spmv + n * s for it in iters)


def make_workload(args, torch, dtype=None):
    from pict_amd import PISOtorch_simulation as S
    from pict_amd import configs
    dtype = dtype or torch.float32
    if args.config == "c5":
        dom, dt, _ = configs.channel_c5(dtype=dtype, device="cuda")
        workload = ("3-D channel 192x96x96 (tanh wall-normal stretch 2.0, periodic x/z, no-slip walls), Re_tau 180, "
                    "FP32, static forcing + random-init 3-layer Conv3d SGS forcing (16 channels) recomputed every "
                    "step, %s" % ("%d-step rollout with per-step checkpointing, MSE loss, forward+backward"
                                  % args.steps if args.mode == "fwdbwd" else "forward PISO steps"))
        sim = S.Simulation(dom, time_step=dt, corrector_steps=2, non_orthogonal=False, advection_tol=1e-6,
                           pressure_tol=1e-6, advect_passive_scalar=False)
        sim.defer_solver_checks = not args.sync_checks
        return dom, sim, dt, workload
    if args.config == "c3":
        dom, dt = configs.periodic_box_c3(256, dtype=dtype, device="cuda")
        workload = ("periodic box 256x256 (2 blocks of 256x128), FP32, spectral k^-5/3 init, learnable per-cell "
                    "forcing, %s" % ("%d-step rollout with per-step checkpointing, MSE loss, forward+backward"
                                     % args.steps if args.mode == "fwdbwd" else "forward PISO steps"))
    else:
        dom = configs.lid_cavity(32, viscosity=0.01, dtype=dtype, device="cuda")
        dt = 0.01
        workload = "lid-driven cavity 32x32, Re=100, FP32, " + ("forward+backward" if args.mode == "fwdbwd"
                                                                else "forward PISO steps")
    for b in dom.blocks:
        b.setVelocitySource(torch.zeros_like(b.velocity))
    dom.UpdateDomainData()
    sim = S.Simulation(dom, time_step=dt, corrector_steps=2, non_orthogonal=False, advection_tol=1e-6,
                       pressure_tol=1e-6, advect_passive_scalar=False)
    sim.defer_solver_checks = not args.sync_checks
    return dom, sim, dt, workload


def target_state(dom, args, torch):
    from pict_amd import configs
    if args.config == "c3":
        vel = configs.spectral_velocity_2d(256, seed=1)
        half = 128
        return [vel[None, :, i * half:(i + 1) * half, :].to(dtype=torch.float32, device="cuda").contiguous()
                for i in range(2)]
    g = torch.Generator().manual_seed(1)
    return [(torch.rand(b.velocity.shape, generator=g) * 0.2).cuda() for b in dom.blocks]


_C5_NET = {}


def rollout_c5(dom, sim, steps, target, torch, keep_tape=False):
    """The 3-D channel's differentiable rollout: before every step the velocity source is the static forcing plus the
    SGS CNN of the current velocity; MSE loss to the target; gradients of the CNN weights."""
    from pict_amd import checkpoint as CK
    from pict_amd import configs
    net = _C5_NET.get("net")
    if net is None:
        net = _C5_NET["net"] = configs.sgs_cnn(16, seed=0)
    blk = dom.blocks[0]
    u0 = blk.velocity.detach().clone()
    forcing = torch.tensor([1.0, 0.0, 0.0], dtype=u0.dtype, device=u0.device).reshape(1, 3, 1, 1, 1)
    net.zero_grad(set_to_none=True)
    out = CK.rollout_with_sources(sim, steps, lambda u: [forcing + 0.05 * net(u[0])], keep_tape=keep_tape)
    loss = ((out[0] - target[0]) ** 2).mean()
    loss.backward()
    sim.flush_solver_checks()
    # back to the initial state (the source keeps its per-cell shape: its gradient tensor has that shape)
    blk.setVelocity(u0)
    blk.setVelocitySource(forcing.expand_as(u0).contiguous())
    dom.UpdateDomainData()
    return [p.grad for p in net.parameters()]


def rollout_fwdbwd(dom, sim, steps, target, torch, keep_tape=False):
    from pict_amd import checkpoint as CK
    u0 = [b.velocity.detach().clone() for b in dom.blocks]
    srcs = [torch.zeros_like(b.velocity).requires_grad_(True) for b in dom.blocks]
    for b, s in zip(dom.blocks, srcs):
        b.setVelocitySource(s)
    dom.UpdateDomainData()
    out = CK.rollout(sim, steps, params=srcs, checkpoint=True, keep_tape=keep_tape)
    loss = sum(((o - t) ** 2).mean() for o, t in zip(out, target))
    loss.backward()
    sim.flush_solver_checks()
    # back to the initial state for the next rollout
    for b, u in zip(dom.blocks, u0):
        b.setVelocity(u)
        b.setVelocitySource(torch.zeros_like(u))
    dom.UpdateDomainData()
    return [s.grad for s in srcs]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--mode", default="fwdbwd", choices=["fwdbwd", "fwd"])
    ap.add_argument("--steps", type=int, default=None, help="default: 16 (c3), 50 (c1), 200 (c2), 100 (c4), 8 (c5)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c3", choices=["c3", "c1", "c2", "c4", "c5"])
    ap.add_argument("--batch", type=int, default=64, help="c4: number of cavities")
    ap.add_argument("--precision-samples", type=int, default=4,
                    help="c3/c1/c5: FP64 states along the trajectory from which the FP32 step's error is measured")
    ap.add_argument("--matched-samples", type=int, default=10,
                    help="c4: FP64 states along the trajectory from which both FP32 paths take one step for the "
                         "per-step error")
    ap.add_argument("--sync-checks", action="store_true",
                    help="c1/c3/c5: check every native step's solver results before queueing the next, as the "
                         "reference's driver does (default: queue without waiting and check them, in order, at the "
                         "end of each rollout)")
    ap.add_argument("--no-batched", action="store_true",
                    help="c3/c1: skip the batching configuration (64 cavities, 100 steps) reported under `batched`")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="CPU baseline sample: whole steps until this much time has passed (at most 64 steps)")
    ap.add_argument("--copy-check", type=int, default=0, metavar="MIB",
                    help="after the timed region, run 3 plain device copies of MIB MiB (a known-size kernel that "
                         "calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE counters)")
    ap.add_argument("--pmc-fetch", default=None, metavar="CSV",
                    help="rocprofv3 --pmc FETCH_SIZE counter_collection.csv of a run with --copy-check 256: fills "
                         "roofline.traffic (with --pmc-write)")
    ap.add_argument("--pmc-write", default=None, metavar="CSV", help="the same for WRITE_SIZE")
    ap.add_argument("--roofline-steps", type=int, default=5,
                    help="c4 / batched: extra steps whose solve launches are timed one by one (HIP events) for the "
                         "pressure CG's roofline")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = {"c1": 50, "c2": 200, "c4": 100, "c5": 8}.get(args.config, 16)
    if args.config == "c5":
        # one FP32 oracle step of 1.77 M cells takes minutes on the host: no CPU sample for this configuration
        args.no_cpu_baseline = True
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the PISO step of one domain runs on a single MI355X "
              "(multi-GPU sharding is a design note, not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from pict_amd import PISOtorch as P

    if args.copy_check > 0:
        # a known-size kernel for the counter calibration (rocprofv3 --pmc runs)
        src = torch.ones(args.copy_check * (1 << 18), dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.config == "c4":
        print(json.dumps(bench_c4(args, torch)))
        return
    if args.config == "c2":
        print(json.dumps(bench_c2(args, torch)))
        return
    dom, sim, dt, workload = make_workload(args, torch)
    N = dom.getTotalSize()
    target = target_state(dom, args, torch)

    def run(k):
        if args.mode == "fwdbwd":
            return (rollout_c5 if args.config == "c5" else rollout_fwdbwd)(dom, sim, k, target, torch)
        assert sim._PISO_split_step(k)
        sim.flush_solver_checks()
        return None

    for _ in range(args.warmup):
        run(2)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record()
    grads = run(args.steps)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = e0.elapsed_time(e1)
    _note("timed region: %d steps, %.3f ms/step" % (args.steps, ms_total / args.steps))
    if grads is not None:
        assert all(torch.isfinite(g).all() for g in grads)
    # the same steps again with the solver log on (HIP events around every solver launch, which add host
    # synchronisations): iteration counts and solver kernel times for the roofline, outside the timed region
    P.enable_solve_log(True)
    run(args.steps)
    torch.cuda.synchronize()
    log = P.solve_log
    P.enable_solve_log(False)
    _note("solver log pass done")

    # per-kernel roofline from the solver log
    agg = {}
    for kind, n, nnz, nrhs, s, infos, kms, transposed, name in log:
        a = agg.setdefault(name, {"ms": 0.0, "bytes": 0, "launches": 0, "iters": 0, "adjoint_launches": 0})
        its = [i.usedIterations + 1 for i in infos]
        a["ms"] += kms
        a["bytes"] += solver_bytes(kind, n, nnz, nrhs, s, its)
        a["launches"] += 1
        a["adjoint_launches"] += int(transposed)
        a["iters"] += sum(its)
    dom_kernel = max(agg, key=lambda k: agg[k]["ms"])
    d = agg[dom_kernel]
    avg_ms = d["ms"] / d["launches"]
    bytes_per_launch = d["bytes"] / d["launches"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    ms_per_step = ms_total / args.steps
    value = N * args.steps / (ms_total * 1e-3)

    extra = {}
    if args.mode == "fwdbwd":
        # the same rollout keeping each step's tape (no recomputation), and the forward alone, for context
        roll = rollout_c5 if args.config == "c5" else rollout_fwdbwd
        for label, fn in (("keep_tape", lambda: roll(dom, sim, args.steps, target, torch, True)),
                          ("forward_only", lambda: (sim._PISO_split_step(args.steps), sim.flush_solver_checks()))):
            torch.cuda.synchronize()
            f0, f1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            f0.record()
            fn()
            f1.record()
            torch.cuda.synchronize()
            extra[label + "_cells_steps_per_s"] = N * args.steps / (f0.elapsed_time(f1) * 1e-3)


    traffic, traffic_source = None, None
    if args.pmc_fetch and args.pmc_write:
        traffic, traffic_source = pmc_traffic(args.pmc_fetch, args.pmc_write, dom_kernel), "counter files"
    elif args.mode == "fwdbwd" and args.config in ("c3", "c5"):
        prof = os.path.join(ROOT, "profiles", "pict_piso_step_fwdbwd_periodic_box_256x256.md" if args.config == "c3"
                            else "pict_piso_step_fwdbwd_channel_192x96x96.md")
        traffic = profiled_traffic(prof, dom_kernel)
        traffic_source = os.path.relpath(prof, ROOT) if traffic is not None else None

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(dom, dt, args, target)
        _note("CPU baseline done")

    # FP32 error of the forward steps against FP64 steps of the same workload (outside the timed region)
    precision = precision_vs_fp64(args, torch, args.steps, args.precision_samples)
    _note("precision against FP64 done")

    batched = None
    if not args.no_batched and args.config == "c3":
        import copy
        bargs = copy.copy(args)
        bargs.steps, bargs.warmup, bargs.no_cpu_baseline = 100, 2, True
        r = bench_c4(bargs, torch)
        c4 = r["config"]
        batched = {"metric": r["metric"], "value": r["value"], "unit": r["unit"], "ms_per_step": r["ms_per_step"],
                   "steps": bargs.steps, "cells": c4["cells"], "workload": c4["workload"],
                   "plain_loop": c4["plain_loop"], "batched_vs_plain": c4["batched_vs_plain"],
                   "error_vs_fp64": c4["error_vs_fp64"], "tuned": c4["tuned"],
                   "solver_ms_per_step": c4["solver_ms_per_step"], "roofline": r["roofline"]}

    out = {
        "metric": "cells*PISO-steps/s (%s)" % ("forward+backward" if args.mode == "fwdbwd" else "forward"),
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
        "config": {"workload": workload, "cells": N, "config": args.config, "mode": args.mode, "dt": dt,
                   "error_vs_fp64": precision,
                   "solver_checks": "after every step" if args.sync_checks else "deferred to the end of the rollout",
                   "solver_launches": {k: v["launches"] for k, v in agg.items()},
                   "solver_adjoint_launches": {k: v["adjoint_launches"] for k, v in agg.items()},
                   "solver_iterations_per_step": {k: v["iters"] / args.steps for k, v in agg.items()},
                   "solver_ms_per_step": {k: v["ms"] / args.steps for k, v in agg.items()},
                   "wall_s": wall, **extra},
        "roofline": {"bound": "hbm", "kernel": dom_kernel, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_source, "bytes_per_launch": bytes_per_launch, "avg_launch_ms": avg_ms},
        "cpu_baseline": cpu,
    }
    if batched is not None:
        out["batched"] = batched
    print(json.dumps(out))


def _c2_sim(torch, dtype, refine):
    from pict_amd import PISOtorch_simulation as S
    from pict_amd import configs
    dom, out, vin = configs.cylinder_ogrid_c2(dtype=dtype, device="cuda")
    sim = S.Simulation(dom, time_step=0.005, corrector_steps=2, advection_tol=1e-6, pressure_tol=1e-6,
                       advect_passive_scalar=False, pressure_refine_tol=refine,
                       prep_fn={"PRE": lambda domain, time_step, **kw: S.update_advective_boundaries(
                           domain, out, vin, time_step.cuda())})
    return dom, sim


def _max_rel(doms, refs, field):
    e = 0.0
    for d, r in zip(doms, refs):
        for b, rb in zip(d.blocks, r.blocks):
            a, c = getattr(b, field).double(), getattr(rb, field).double()
            e = max(e, float((a - c).abs().max() / c.abs().max()))
    return e


def bench_c2(args, torch):
    """Precision configuration: the 4-block O-grid cylinder (4 x 128 x 128 cells, inflow u = 1, advective outflow,
    nu = 0.005, noisy uniform initial state) stepped with the reference driver's defaults (non-orthogonal branch,
    flags 25), dt = 0.005, advection and pressure tolerance 1e-6, FP32 bulk with FP64 Krylov reductions and the
    pressure solves refined in FP64 to 1e-10 relative residual.  Times `--steps` steps after `--warmup`; then runs
    the same steps in FP64 (FP64 refined pressure solves) and in plain FP32 (no refinement) from the same initial
    state and reports the error of both FP32 runs against the FP64 one."""
    from pict_amd import PISOtorch as P
    dom, sim = _c2_sim(torch, torch.float32, 1e-10)
    N = dom.getTotalSize()
    sim.record_solver_infos = True
    if args.warmup > 0:
        assert sim._PISO_split_step(args.warmup)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    sim.solver_infos = []
    t0 = time.perf_counter()
    e0.record()
    assert sim._PISO_split_step(args.steps)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = e0.elapsed_time(e1)
    _note("c2: timed region %d steps, %.1f ms/step" % (args.steps, ms_total / args.steps))
    infos = sim.solver_infos
    adv_its = [sum(i.usedIterations for i in inf) for k, inf in infos if k == "BiCG"]
    p_its = [inf[0].usedIterations for k, inf in infos if k == "CG"]
    # FP64 (refined) and plain FP32 runs of the same steps from the same initial state
    ref, sref = _c2_sim(torch, torch.float64, 1e-10)
    assert sref._PISO_split_step(args.warmup + args.steps)
    _note("c2: FP64 run done")
    plain, splain = _c2_sim(torch, torch.float32, None)
    assert splain._PISO_split_step(args.warmup + args.steps)
    _note("c2: plain FP32 run done")
    err = {"refined": {"velocity": _max_rel([dom], [ref], "velocity"),
                       "pressure": _max_rel([dom], [ref], "pressure")},
           "plain_fp32": {"velocity": _max_rel([plain], [ref], "velocity"),
                          "pressure": _max_rel([plain], [ref], "pressure")},
           "reference": "FP64 steps of the same path with FP64 refined pressure solves (each step matches the FP64 "
                        "oracle's non-orthogonal step to 1e-8 at reduced size, tests/test_gpu_c2.py)"}
    # solver kernel times of two more steps (HIP events around every launch) for the roofline
    P.enable_solve_log(True)
    assert sim._PISO_split_step(2)
    torch.cuda.synchronize()
    log = P.solve_log
    P.enable_solve_log(False)
    agg = {}
    for kind, n, nnz, nrhs, s_, inf, kms, transposed, name in log:
        a = agg.setdefault(name, {"ms": 0.0, "bytes": 0, "launches": 0, "iters": 0})
        its = [i.usedIterations + 1 for i in inf]
        a["ms"] += kms
        a["bytes"] += solver_bytes(kind, n, nnz, nrhs, s_, its)
        a["launches"] += 1
        a["iters"] += sum(its)
    dk = max(agg, key=lambda k: agg[k]["ms"])
    d = agg[dk]
    avg_ms = d["ms"] / d["launches"]
    bpl = d["bytes"] / d["launches"]
    achieved = bpl / (avg_ms * 1e-3) / 1e9
    cpu = None if args.no_cpu_baseline else cpu_baseline_c2(args)
    return {
        "metric": "cells*PISO-steps/s (forward)",
        "value": N * args.steps / (ms_total * 1e-3),
        "unit": "cells*steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 Krylov reductions, FP64-refined pressure solves)",
        "data": "synthetic (seeded)",
        "config": {"workload": "O-grid around a cylinder r=0.5, 4 blocks of 128x128 (angle x radius, tanh radial "
                               "stretch, first ratio 1.05, outer radius 5), inflow u=1 and advective outflow, "
                               "nu=0.005, u0 = (1,0) + N(0, 0.01^2), dt=0.005, reference driver defaults "
                               "(non-orthogonal branch, flags 25), tolerances 1e-6, pressure refined to 1e-10",
                   "cells": N, "config": "c2", "mode": "fwd", "dt": 0.005, "wall_s": wall,
                   "error_vs_fp64": err,
                   "iterations_per_step": {"advection": sum(adv_its) / args.steps,
                                           "pressure_inner": sum(p_its) / args.steps},
                   "solver_ms_per_step": {k: v["ms"] / 2 for k, v in agg.items()},
                   "solver_launches_per_step": {k: v["launches"] / 2 for k, v in agg.items()}},
        "roofline": {"bound": "hbm", "kernel": dk, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "bytes_per_launch": bpl,
                     "avg_launch_ms": avg_ms},
        "cpu_baseline": cpu,
    }


def _cpu_threads():
    """OpenMP threads of the CPU sample: the cores this process may run on, at most the host's CPU share for one
    GPU (16 threads, or OMP_NUM_THREADS when it is set lower)."""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        share = int(os.environ.get("OMP_NUM_THREADS", "16"))
    except ValueError:
        share = 16
    return max(1, min(avail, share, 16))


def _cpu_host():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    return {"model": model, "logical_cpus": os.cpu_count(), "cpus_available": avail}


def _cpu_sample(step, seconds, threads, max_steps=64):
    """step() (returns the cells it stepped, and the seconds to count: None = all of it) on `threads` OpenMP threads
    of the oracle until `seconds` have passed, at least once.  Returns (cells*steps/s, steps, seconds)."""
    from oracle import oracle as O
    O.lib(32).oracle_set_threads(threads)
    t = 0.0
    cells = steps = 0
    while steps < max_steps and (steps == 0 or t < seconds):
        t1 = time.perf_counter()
        c, dt_counted = step(steps)
        t += dt_counted if dt_counted is not None else time.perf_counter() - t1
        cells += c
        steps += 1
    return cells / t, steps, t


def _cpu_result(step, args, what):
    """The CPU baseline entry: the same sample on the host's CPU share and on one thread."""
    threads = _cpu_threads()
    v, n, t = _cpu_sample(step, args.cpu_seconds, threads)
    v1, n1, t1 = _cpu_sample(step, args.cpu_seconds, 1)
    return {"value": v, "unit": "cells*steps/s", "cores": threads, "kind": "port",
            "sample": "%d %s in %.1f s (FP32 C restatement, OpenMP %d threads)" % (n, what, t, threads),
            "single_thread": {"value": v1, "sample": "%d %s in %.1f s, 1 thread" % (n1, what, t1)},
            "host": _cpu_host()}


def cpu_baseline_c2(args):
    """FP32 C restatement (OpenMP) of the precision configuration's non-orthogonal step (plain FP32 solves at the
    same tolerances; the outflow update on the host twin), until --cpu-seconds have passed."""
    import torch
    from oracle import oracle as O
    from pict_amd import PISOtorch_simulation as S
    from pict_amd import configs
    dom, out, vin = configs.cylinder_ogrid_c2(dtype=torch.float32, device="cuda")
    twin = O.cpu_twin(dom, dtype=torch.float32)
    tout = [twin.getBlock(i).getBoundary("+y") for i in (0, 1)]
    vin = vin.cpu()
    dt = torch.tensor([0.005], dtype=torch.float32)

    def step(_):
        S.update_advective_boundaries(twin, tout, vin, dt)
        O.piso_step(twin, 0.005, corrector_steps=2, advection_tol=1e-6, pressure_tol=1e-6, non_orthogonal=True)
        return dom.getTotalSize(), None
    return _cpu_result(step, args, "non-orthogonal PISO step(s) of the same workload from its initial state "
                                   "onwards (FP32 without refinement)")


def _c4_sims(torch, count, dtype):
    from pict_amd import PISOtorch_simulation as S
    from pict_amd import configs
    doms, ns, visc, dt = configs.cavity_batch_c4(count, 0, dtype, "cuda")
    # no preconditioned retry of a failed BiCGStab in either leg: the batched solver has none, and the plain loop
    # would otherwise keep each step's input velocity for it
    sims = [S.Simulation(d, time_step=dt, corrector_steps=2, non_orthogonal=False, advection_tol=1e-6,
                         pressure_tol=1e-6, advect_passive_scalar=False, BiCG_precondition_fallback=False)
            for d in doms]
    return doms, sims, ns, visc, dt


def _c4_error(doms, ref):
    """Max over cavities of max|u - u_ref| / max|u_ref| (and the same for the pressure)."""
    ev = ep = 0.0
    for d, r in zip(doms, ref):
        for b, rb in zip(d.blocks, r.blocks):
            ev = max(ev, float((b.velocity.double() - rb.velocity.double()).abs().max() /
                               rb.velocity.double().abs().max()))
            ep = max(ep, float((b.pressure.double() - rb.pressure.double()).abs().max() /
                               rb.pressure.double().abs().max()))
    return ev, ep


def _fixed_count_sim(torch, dom, dt):
    """Stage-path simulation whose solves each run a given number of iterations (tolerance 0, list `counts`, one
    entry per solve and right-hand side in the driver's order): the reference step of a precision comparison takes
    the iteration counts of the FP32 step it is compared with, so that the comparison measures rounding, not on
    which side of the stopping tolerance a solve lands (one CG iteration near the tolerance moves a cavity's
    velocity by about 1e-3 relative)."""
    from pict_amd import PISOtorch as P
    from pict_amd import PISOtorch_simulation as S

    class FixedCountSim(S.Simulation):
        counts = ()

        def linear_solve(self, csrMat, rhs, x=None, use_BiCG=False, **kw):
            if x is None:
                x = torch.zeros_like(rhs)
            n = csrMat.getRows()
            for k in range(rhs.numel() // n):
                # used iterations as the solvers report them: CG the index of its last iteration (-1: none),
                # BiCGStab its full iterations (one converging at a half step reports the full ones before it)
                it = self.counts.pop(0) + (0 if use_BiCG else 1)
                if it > 0:
                    P.SolveLinear(csrMat, rhs[k * n:(k + 1) * n], x[k * n:(k + 1) * n], maxIterations=it,
                                  tolerance=0.0, useBiCG=use_BiCG, BiCGwithPreconditioner=False)
            return x, True

    s = FixedCountSim(dom, time_step=dt, corrector_steps=2, non_orthogonal=False, advection_tol=1e-6,
                      pressure_tol=1e-6, advect_passive_scalar=False)
    s.native_step = False
    return s


def _snapshot(dom):
    return [(b.velocity.detach().clone(), b.pressure.detach().clone()) for b in dom.blocks]


def _set_state(dom, snap):
    """Velocity and pressure of every block from `snap`, copied in place (cast to the domain's dtype; the tensors
    keep their addresses)."""
    for b, (u, p) in zip(dom.blocks, snap):
        b.velocity.detach().copy_(u)
        b.pressure.detach().copy_(p)
    dom.UpdateDomainData()


def _dom_error(dom, ref):
    """max|u - u_ref| / max|u_ref| over all blocks, and the same for the pressure."""
    out = []
    for f in ("velocity", "pressure"):
        diff = max(float((getattr(b, f).detach().double() - getattr(r, f).detach().double()).abs().max())
                   for b, r in zip(dom.blocks, ref.blocks))
        scale = max(float(getattr(r, f).detach().double().abs().max()) for r in ref.blocks)
        out.append(diff / scale)
    return out


def _counted_step(sim):
    """One step of `sim` (FP32); returns the used iterations of its solves in the driver's order."""
    sim.record_solver_infos = True
    sim.solver_infos = []
    assert sim._PISO_split_step(1)
    sim.record_solver_infos = False
    return [i.usedIterations for _, infos in sim.solver_infos for i in infos]


def _ref_error(dom, counts, snap, ref_dom, ref_sim, source=None):
    """Error of `dom`'s state after one FP32 step from the FP64 state `snap` that took the iteration counts
    `counts`, against the FP64 step of ref_sim from `snap` with those counts (`source(dom)` sets the velocity
    source from the state first)."""
    _set_state(ref_dom, snap)
    if source:
        source(ref_dom)
    ref_sim.counts = list(counts)
    assert ref_sim._PISO_split_step(1)
    assert not ref_sim.counts
    return _dom_error(dom, ref_dom)


def precision_vs_fp64(args, torch, steps, samples):
    """FP32 error of this configuration's forward steps against FP64 (the library's FP64 steps, same inputs and
    solver settings): the state after `steps` steps from the same initial state (trajectory), and one step from
    each of `samples` FP64 states along that trajectory against the FP64 step with the FP32 step's iteration
    counts (per step)."""
    d32, s32, dt, _ = make_workload(args, torch, torch.float32)
    d64, s64, _, _ = make_workload(args, torch, torch.float64)
    dm, _, _, _ = make_workload(args, torch, torch.float64)
    sm = _fixed_count_sim(torch, dm, dt)
    s32.defer_solver_checks = s64.defer_solver_checks = False
    source = None
    if args.config == "c5":
        from pict_amd import configs
        nets = {dt_: configs.sgs_cnn(16, seed=0, dtype=dt_) for dt_ in (torch.float32, torch.float64)}

        def source(dom):
            b = dom.blocks[0]
            u = b.velocity.detach()
            f = torch.tensor([1.0, 0.0, 0.0], dtype=u.dtype, device=u.device).reshape(1, 3, 1, 1, 1)
            with torch.no_grad():
                b.setVelocitySource((f + 0.05 * nets[u.dtype](u)).contiguous())
            dom.UpdateDomainData()
    snaps = []
    every = max(1, steps // samples)
    for k in range(steps):
        if k % every == 0:
            snaps.append(_snapshot(d64))
        for d, s in ((d32, s32), (d64, s64)):
            if source:
                source(d)
            assert s._PISO_split_step(1)
    ev, ep = _dom_error(d32, d64)
    per = [0.0, 0.0]
    for snap in snaps:
        _set_state(d32, snap)
        if source:
            source(d32)
        e = _ref_error(d32, _counted_step(s32), snap, dm, sm, source)
        per = [max(per[0], e[0]), max(per[1], e[1])]
    # the same states with the pressure solves refined in FP64 to 1e-10 relative residual in both precisions:
    # what is left is the rounding of the FP32 stages, the advection solve and the FP32 pressure system
    s32.pressure_refine_tol = s64.pressure_refine_tol = 1e-10
    ref = [0.0, 0.0]
    for snap in snaps:
        for d, s in ((d32, s32), (d64, s64)):
            _set_state(d, snap)
            if source:
                source(d)
            assert s._PISO_split_step(1)
        e = _dom_error(d32, d64)
        ref = [max(ref[0], e[0]), max(ref[1], e[1])]
    s32.pressure_refine_tol = s64.pressure_refine_tol = None
    torch.cuda.synchronize()
    return {"trajectory": {"velocity": ev, "pressure": ep, "steps": steps},
            "per_step": {"velocity": per[0], "pressure": per[1], "states": len(snaps)},
            "per_step_refined_pressure": {"velocity": ref[0], "pressure": ref[1]},
            "reference": "FP64 steps of the same workload and solver settings; per step: one step from each FP64 "
                         "state (cast to FP32) against the FP64 step from that state whose solves run the FP32 "
                         "step's iteration counts; refined: both steps with FP64-refined pressure solves"}


def _c4_matched_error(torch, snaps, doms_p, sims_p, batch, doms_b, dt):
    """Per-step precision of both FP32 paths on the same inputs: from each FP64 snapshot (a state per cavity), one
    FP32 step of the plain loop and one batched step against the FP64 step with the iteration counts of that FP32
    step.  Returns {"plain": {"velocity", "pressure"}, "batched": ...}, maxima over snapshots and cavities."""
    from pict_amd import configs
    doms_r, _, _, _ = configs.cavity_batch_c4(len(doms_p), 0, torch.float64, "cuda")
    sims_r = [_fixed_count_sim(torch, d, dt) for d in doms_r]
    out = {"plain": [0.0, 0.0], "batched": [0.0, 0.0]}
    for snap in snaps:
        for d, sn in zip(doms_b, snap):
            _set_state(d, sn)
        counts_b = [[int(c) for c in row] for row in batch.step()["used_iterations"]]
        for i, (dp, sp, db, dr, sr) in enumerate(zip(doms_p, sims_p, doms_b, doms_r, sims_r)):
            _set_state(dp, snap[i])
            res = [_ref_error(dp, _counted_step(sp), snap[i], dr, sr), _ref_error(db, counts_b[i], snap[i], dr, sr)]
            for name, e in zip(("plain", "batched"), res):
                out[name] = [max(out[name][0], e[0]), max(out[name][1], e[1])]
    return {k: {"velocity": v[0], "pressure": v[1]} for k, v in out.items()}


def bench_c4(args, torch):
    """Batching configuration: `--batch` independent lid cavities (sides 32..256, viscosities 1e-3..1e-1, FP32,
    BiCGStab advection and CG pressure at tolerance 1e-6), `--steps` PISO steps each.  Times the batched step
    (SimulationBatch: one launch per stage and per solve, graph replay, autotuned launch shapes) and the plain loop
    over the same simulations (one native step per simulation and step), both from the state after `--warmup`
    steps; both end states are compared with FP64 batched steps of the same cavities and solver settings."""
    import numpy as np
    from pict_amd.batch import SimulationBatch
    doms_b, sims_b, ns, visc, dt = _c4_sims(torch, args.batch, torch.float32)
    doms_p, sims_p, _, _, _ = _c4_sims(torch, args.batch, torch.float32)
    N = sum(n * n for n in ns)
    batch = SimulationBatch(sims_b)
    t_tune = time.perf_counter()
    tuned = batch.tune(repeats=3)
    t_tune = time.perf_counter() - t_tune
    _note("batched: launch shapes tuned in %.1f s" % t_tune)
    for _ in range(args.warmup):
        batch.step()
        for s in sims_p:
            assert s._PISO_split_step(1)
    torch.cuda.synchronize()

    def timed(fn):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    iters = []

    def run_batched():
        for _ in range(args.steps):
            iters.append(batch.step()["used_iterations"].copy())

    def run_plain():
        for _ in range(args.steps):
            for s in sims_p:
                assert s._PISO_split_step(1)

    ms_b = timed(run_batched)
    _note("batched: %d timed steps, %.3f ms/step" % (args.steps, ms_b / args.steps))
    ms_p = timed(run_plain)
    _note("batched: plain loop %.3f ms/step" % (ms_p / args.steps))
    info = batch.describe()
    # FP64 steps of the same cavities and settings: the error of both FP32 paths
    doms_r, sims_r, _, _, _ = _c4_sims(torch, args.batch, torch.float64)
    ref = SimulationBatch(sims_r)
    snaps = []
    total = args.warmup + args.steps
    for k in range(total):
        if k % max(1, total // args.matched_samples) == 0:
            snaps.append([_snapshot(d) for d in doms_r])
        ref.step()
    _note("batched: FP64 reference steps done")
    ev_b, ep_b = _c4_error(doms_b, doms_r)
    ev_p, ep_p = _c4_error(doms_p, doms_r)
    # roofline of the batched pressure CG, the dominant kernel: `--roofline-steps` more batched steps with events
    # around each solve launch (no graph), each launch's algorithmic bytes from its own iteration counts
    K = 2

    def launch_bytes(used, k):
        return sum(solver_bytes("CG", d.getTotalSize(), d.C.getSize(), 1, 4, [int(used[i][2 + k]) + 1])
                   for i, d in enumerate(doms_b))
    batch.set_timing(True)
    l_ms, l_bytes, kms_all = [], [], []
    for _ in range(max(1, args.roofline_steps)):
        arr = batch.step()
        kms = batch.last_solve_ms()
        kms_all.append(kms)
        for k in range(K):
            l_ms.append(kms[1 + k])
            l_bytes.append(launch_bytes(arr["used_iterations"], k))
    batch.set_timing(False)
    kms = [float(np.mean([x[j] for x in kms_all])) for j in range(len(kms_all[0]))]
    avg_ms = float(np.mean(l_ms))
    bytes_per_launch = float(np.mean(l_bytes))
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    # the same model over the timed (graph-replayed) steps, for a kernel trace of them
    timed_bytes_per_launch = float(np.mean([launch_bytes(u, k) for u in iters for k in range(K)]))
    traffic, traffic_source = None, None
    if args.pmc_fetch and args.pmc_write:
        traffic, traffic_source = pmc_traffic(args.pmc_fetch, args.pmc_write, info["pressure_kernel"]), "counter files"
    else:
        prof = os.path.join(ROOT, "profiles", "pict_batch_step_64_cavities.md")
        traffic = profiled_traffic(prof, info["pressure_kernel"])
        traffic_source = os.path.relpath(prof, ROOT) if traffic is not None else None
    its = np.array(iters)
    _note("batched: roofline launches timed")
    matched = _c4_matched_error(torch, snaps, doms_p, sims_p, batch, doms_b, dt)
    _note("batched: per-step precision done")
    cpu = None if args.no_cpu_baseline else cpu_baseline_c4(doms_b, dt, args)
    value = N * args.steps / (ms_b * 1e-3)
    plain = N * args.steps / (ms_p * 1e-3)
    return {
        "metric": "cells*PISO-steps/s (forward, batched %d cavities)" % args.batch,
        "value": value,
        "unit": "cells*steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_b / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 Krylov reductions)",
        "data": "synthetic (seeded)",
        "config": {"workload": "%d lid-driven cavities, sides %s (uniform over 32..256, seed 0), viscosities "
                               "log-uniform 1e-3..1e-1, FP32, dt %.4g, BiCGStab/CG tolerance 1e-6, 2 correctors"
                               % (args.batch, sorted(set(ns)), dt),
                   "cells": N, "config": "c4", "mode": "fwd",
                   "batched_vs_plain": value / plain,
                   "plain_loop": {"value": plain, "ms_per_step": ms_p / args.steps},
                   "error_vs_fp64": {"per_step": dict(matched, reference=(
                                         "one step from each of %d FP64 states along the trajectory (the state cast "
                                         "to FP32) against the FP64 step from that state whose solves run the FP32 "
                                         "step's iteration counts; max over states and cavities" % len(snaps))),
                                     "trajectory": {"batched": {"velocity": ev_b, "pressure": ep_b},
                                                    "plain": {"velocity": ev_p, "pressure": ep_p},
                                                    "reference": "FP64 batched steps, same cavities and solver "
                                                                 "settings (tolerance 1e-6 stopping rule)"}},
                   "tuned": tuned, "tune_s": t_tune, "launch": info,
                   "iterations_per_step": {"advection_max": int(its[:, :, :2].max()),
                                           "pressure_max": int(its[:, :, 2:].max()),
                                           "pressure_sum": float(its[:, :, 2:].sum() / args.steps)},
                   "solver_ms_per_step": {"advection": kms[0], "pressure": sum(kms[1:])}},
        "roofline": {"bound": "hbm", "kernel": info["pressure_kernel"], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_source, "bytes_per_launch": bytes_per_launch, "avg_launch_ms": avg_ms,
                     "launches": len(l_ms), "timed_bytes_per_launch": timed_bytes_per_launch},
        "cpu_baseline": cpu,
    }


def cpu_baseline_c4(doms, dt, args):
    """FP32 C restatement (OpenMP) of one forward step of each cavity in turn, until --cpu-seconds have passed."""
    import torch
    from oracle import oracle as O

    def step(i):
        d = doms[i % len(doms)]
        twin = O.cpu_twin(d, dtype=torch.float32)  # set-up outside the timed part
        t1 = time.perf_counter()
        O.piso_step(twin, dt, corrector_steps=2, advection_tol=1e-6, pressure_tol=1e-6)
        return d.getTotalSize(), time.perf_counter() - t1
    return _cpu_result(step, args, "PISO steps, one of each cavity of the same batch in turn from the first")


def pmc_traffic(fetch_csv, write_csv, kernel):
    """HBM bytes per launch of `kernel` from rocprofv3 FETCH_SIZE / WRITE_SIZE counter files (KiB per dispatch),
    each scaled by what the same counter reads for the 256 MiB device copy of --copy-check 256."""
    import csv

    def per_kernel(path, counter):
        vals, copy = [], 0.0
        with open(path) as fh:
            for r in csv.DictReader(fh):
                if r["Counter_Name"] != counter:
                    continue
                name = r["Kernel_Name"].replace("void ", "").split("(")[0]
                v = float(r["Counter_Value"])
                if name == kernel:
                    vals.append(v)
                elif "copyBuffer" in name:
                    copy = max(copy, v)
        if not vals or copy <= 0:
            return None
        return sum(vals) / len(vals) * (256 * 1024 / copy) * 1024

    f, w = per_kernel(fetch_csv, "FETCH_SIZE"), per_kernel(write_csv, "WRITE_SIZE")
    return None if f is None or w is None else f + w


def profiled_traffic(path, kernel):
    """HBM bytes per launch of `kernel` as recorded in a committed profile summary: the 'scaled HBM MB per launch'
    column of its FETCH_SIZE / WRITE_SIZE table (same workload, rocprofv3 --pmc runs); None if absent."""
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            cols = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cols) == 6 and cols[0] == "`%s`" % kernel:
                try:
                    return float(cols[4]) * 1e6
                except ValueError:
                    return None
    return None


def cpu_baseline(dom, dt, args, target):
    """FP32 C restatement of the same step (oracle/pict_oracle.c built with -fopenmp) on a sample of steps: the
    forward step, and for fwdbwd also the recomputed forward and the reverse pass of a checkpointed step."""
    import torch
    from oracle import oracle as O
    twin = O.cpu_twin(dom, dtype=torch.float32)
    g = torch.Generator().manual_seed(0)

    def step(_):
        u_in = [b.velocity.clone() for b in twin.blocks]
        O.piso_step(twin, dt, corrector_steps=2, advection_tol=1e-6, pressure_tol=1e-6)
        if args.mode == "fwdbwd":
            gout = [torch.rand(b.velocity.shape, generator=g) * 1e-3 for b in twin.blocks]
            out = [b.velocity.clone() for b in twin.blocks]
            for b, u in zip(twin.blocks, u_in):
                b.setVelocity(u)
            O.piso_step_vjp(twin, dt, gout, corrector_steps=2, advection_tol=1e-6, pressure_tol=1e-6)
            for b, u in zip(twin.blocks, out):
                b.setVelocity(u)
        return dom.getTotalSize(), None
    what = "forward step(s) + recomputed forward + reverse pass" if args.mode == "fwdbwd" else "forward step(s)"
    return _cpu_result(step, args, "PISO %s of the same workload" % what)


if __name__ == "__main__":
    main()
