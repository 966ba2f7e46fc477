"""Batched PISO steps (pict_amd.batch.SimulationBatch over pict_batch_*): independent simulations of different
sizes, structures, viscosities and boundaries stepped with one launch per stage and one launch per solve.

- Launch shapes that give every system the workgroup count of its own native step: the state after several
  batched steps is bitwise the state after the same native steps run one simulation at a time (the batch runs
  the same cell bodies and solver bodies, and its per-domain reductions keep the single-domain order).
- FP32 lid cavities, default and autotuned shapes: error against the FP64 oracle's steps (CPU, the same solver
  settings: tolerance 1e-6 on the RMS residual) within 1e-5 of the field scale, like the plain per-simulation
  loop's.  (At that tolerance the solves stop well before full convergence, so the reference for the FP32
  error is the FP64 step with the same stopping rule, not the converged solution.)
- The register-resident BiCGStab (rows and vectors in registers, halos through published tagged words) agrees
  with the streaming BiCGStab kernel on systems split over many workgroups, periodic wraps and rotated block
  connections, for 1, 2 and 4 rows per thread.
- From the second step on a batch replays a captured graph: bitwise the steps queued directly.
- The 64-cavity workload (sides 32..256, viscosities 1e-3..1e-1): batched and plain loop agree, every solve
  converges, and the schedule fits the device.
"""
import numpy as np
import pytest
import torch

import scenarios
from oracle import oracle as O
from pict_amd import PISOtorch_simulation as S
from pict_amd import configs
from pict_amd.batch import SimulationBatch

pytestmark = pytest.mark.gpu


def _sim(d, dtype, dt=0.01):
    tol = 1e-10 if dtype == torch.float64 else 1e-6
    return S.Simulation(d, time_step=dt, corrector_steps=2, non_orthogonal=False, advection_tol=tol,
                        pressure_tol=tol, advect_passive_scalar=False)


def _domains(dtype):
    """Lid cavities of several sizes and viscosities (one of 40x40 cells: more than one workgroup per system),
    random-state scenarios with a periodic 2-block box and a rotated block connection with per-cell viscosity."""
    doms = [configs.lid_cavity(n, viscosity=v, dtype=dtype, device="cuda", name="c%d" % i)
            for i, (n, v) in enumerate([(8, 0.05), (12, 0.01), (16, 0.002), (24, 0.02), (33, 0.005), (40, 0.001)])]
    doms.append(scenarios.cavity(dtype=dtype, device="cuda", n=10))
    doms.append(scenarios.periodic_two_block(dtype=dtype, device="cuda"))
    doms.append(scenarios.rotated_connection(dtype=dtype, device="cuda"))
    return doms


def _state(doms):
    return [[b.velocity.detach().clone() for b in d.blocks] + [b.pressure.detach().clone() for b in d.blocks]
            for d in doms]


def _plain(dtype, steps):
    doms = _domains(dtype)
    sims = [_sim(d, dtype) for d in doms]
    its = []
    for s in sims:
        s.record_solver_infos = True
        row = []
        for _ in range(steps):
            assert s._PISO_split_step(1)
            row.append([i.usedIterations for _, inf in s.solver_infos for i in inf])
        its.append(row)
    return doms, sims, its


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_batch_matches_native_steps_bitwise(dtype):
    print("case", dtype, flush=True)
    steps = 3
    ref, _, its_ref = _plain(dtype, steps)
    doms = _domains(dtype)
    sims = [_sim(d, dtype) for d in doms]
    # the single-system plans of these sizes: register-resident CG with one row per thread; BiCGStab register-resident
    # with one row per thread in FP32, streaming with 512 rows per workgroup in FP64
    adv = dict(advection_rows_per_thread=1) if dtype == torch.float32 else dict(advection_rows_per_thread=-1,
                                                                                   advection_rows_per_workgroup=512)
    batch = SimulationBatch(sims, config=dict(pressure_rows_per_thread=1, **adv))
    its = [[] for _ in sims]
    for _ in range(steps):
        arr = batch.step()
        for i in range(len(sims)):
            its[i].append(arr["used_iterations"][i].tolist())
    info = batch.describe()
    print(" batch:", info, flush=True)
    assert info["pressure_rows_per_thread"] == 1
    assert its == its_ref
    for a, b in zip(_state(ref), _state(doms)):
        for x, y in zip(a, b):
            assert torch.equal(x, y)
    for s in sims:
        assert s.total_step == steps


def _cavities(dtype, device="cuda"):
    return [configs.lid_cavity(n, viscosity=v, dtype=dtype, device=device, name="c%d" % i)
            for i, (n, v) in enumerate([(16, 0.01), (24, 0.003), (32, 0.001), (48, 0.02)])]


def _oracle_steps(doms, steps, dt, tol):
    out = []
    for d in doms:
        t = O.cpu_twin(d, torch.float64)
        for _ in range(steps):
            O.piso_step(t, dt, corrector_steps=2, advection_tol=tol, pressure_tol=tol)
        out.append(t)
    return out


def _err(doms, refs):
    ev, ep = 0.0, 0.0
    for d, r in zip(doms, refs):
        for b, rb in zip(d.blocks, r.blocks):
            ev = max(ev, float((b.velocity.double().cpu() - rb.velocity).abs().max()) /
                     max(float(rb.velocity.abs().max()), 1e-30))
            ep = max(ep, float((b.pressure.double().cpu() - rb.pressure).abs().max()) /
                     max(float(rb.pressure.abs().max()), 1e-30))
    return ev, ep


@pytest.mark.parametrize("tuned", [False, True])
def test_batch_fp32_error_against_fp64_oracle(tuned):
    print("case tuned", tuned, flush=True)
    steps, dt = 4, 0.01
    refs = _oracle_steps(_cavities(torch.float64, "cuda"), steps, np.float32(dt).item(), 1e-6)
    plain = _cavities(torch.float32)
    for d in plain:
        s = _sim(d, torch.float32, dt)
        for _ in range(steps):
            assert s._PISO_split_step(1)
    doms = _cavities(torch.float32)
    batch = SimulationBatch([_sim(d, torch.float32, dt) for d in doms])
    if tuned:
        cfg = batch.tune(repeats=2)
        print(" tuned:", cfg, flush=True)
        assert cfg["tiles_per_workgroup"] in (1, 2, 4)
        assert cfg["pressure_rows_per_thread"] in (1, 2, 4, 8)
        assert cfg["advection_rows_per_thread"] in (-1, 1, 2, 4)
        if cfg["advection_rows_per_thread"] == -1:
            assert cfg["advection_rows_per_workgroup"] in (512, 1024, 2048)
        assert batch.config() == cfg
    batch.run(steps)
    ev_b, ep_b = _err(doms, refs)
    ev_p, ep_p = _err(plain, refs)
    print(" FP64-oracle error: batched velocity %.2e pressure %.2e, plain velocity %.2e pressure %.2e"
          % (ev_b, ep_b, ev_p, ep_p), flush=True)
    assert ev_b < 1e-5 and ep_b < 1e-5
    assert ev_p < 1e-5 and ep_p < 1e-5


def _halo_domains(dtype):
    """Systems split over several workgroups at every rows-per-thread shape: larger periodic 2-block boxes (columns
    that wrap around, rows with more foreign entries than the register list holds), a rotated connection, lid
    cavities."""
    doms = [scenarios.periodic_two_block(dtype=dtype, device="cuda", nx=48, ny=20, seed=3),
            scenarios.periodic_two_block(dtype=dtype, device="cuda", nx=130, ny=40, seed=4),
            scenarios.rotated_connection(dtype=dtype, device="cuda")]
    doms += [configs.lid_cavity(n, viscosity=v, dtype=dtype, device="cuda", name="c%d" % i)
             for i, (n, v) in enumerate([(40, 0.08), (72, 0.01)])]
    return doms


@pytest.mark.parametrize("rpt", [1, 2, 4])
def test_batch_register_bicgstab_matches_streaming(rpt):
    print("case rows per thread", rpt, flush=True)
    steps = 3
    out = []
    for cfg in (dict(advection_rows_per_thread=-1, advection_rows_per_workgroup=256),
                dict(advection_rows_per_thread=rpt)):
        doms = _halo_domains(torch.float32)
        batch = SimulationBatch([_sim(d, torch.float32, 0.02) for d in doms], config=cfg)
        its = [batch.step()["used_iterations"].copy() for _ in range(steps)]
        info = batch.describe()
        print(" ", info["advection_kernel"], "grid", info["advection_grid"], "jobs", info["advection_jobs"],
              "advection iterations", [i[:, :2].tolist() for i in its], flush=True)
        out.append((_state(doms), its, info))
    (sa, ia, _), (sb, ib, info) = out
    assert info["advection_kernel"] == "pict::k_bicg_local_jobs<%d, 5, 2>" % rpt
    assert info["advection_jobs"] > len(sb)  # systems spread over several workgroups
    # the two kernels sum a row's entries in different orders: on the stiffest system (the 40x40 cavity with
    # viscosity 0.08) FP32 BiCGStab takes a few iterations more or fewer; both converge
    for x, y in zip(ia, ib):
        assert (np.abs(x[:, :2] - y[:, :2]) <= np.maximum(3, 0.15 * np.abs(x[:, :2]))).all()
    worst = 0.0
    for a, b in zip(sa, sb):
        for x, y in zip(a, b):
            worst = max(worst, float((x - y).abs().max()) / max(float(x.abs().max()), 1e-30))
    print(" register vs streaming BiCGStab: %.2e" % worst, flush=True)
    assert worst < 1e-4


def test_batch_fp64_uses_streaming_bicgstab():
    doms = _domains(torch.float64)
    batch = SimulationBatch([_sim(d, torch.float64) for d in doms])
    batch.step()
    assert batch.describe()["advection_kernel"] == "pict::k_bicgstab_jobs<double, 2>"


def test_batch_graph_replays_match_direct(monkeypatch):
    steps = 5
    out = []
    for graphs in ("0", "1"):
        monkeypatch.setenv("PICT_STEP_GRAPHS", graphs)
        doms = _domains(torch.float32)
        batch = SimulationBatch([_sim(d, torch.float32) for d in doms])
        states = []
        for i in range(steps):
            batch.step()
            # eager work on the same stream between replays
            noise = torch.ones(4096, device="cuda").cumsum(0)
            torch.cuda.synchronize()
            assert float(noise[-1]) == 4096.0
            states.append(_state(doms))
            assert batch.describe()["graph"] == (1 if graphs == "1" and i >= 1 else 0)
        out.append(states)
    for a, b in zip(*out):
        for x, y in zip(a, b):
            for t, u in zip(x, y):
                assert torch.equal(t, u)


def test_batch_tensors_swapped_between_steps():
    """A domain whose tensors are replaced between steps: the tables are rewritten and the replayed graph reads
    the new tensors."""
    doms = _domains(torch.float32)
    ref = _domains(torch.float32)
    batch = SimulationBatch([_sim(d, torch.float32) for d in doms])
    sims_ref = [_sim(d, torch.float32) for d in ref]
    batch.set_config(pressure_rows_per_thread=1, advection_rows_per_thread=1)
    for k in range(4):
        if k == 2:
            for d in (doms[0], ref[0]):
                d.blocks[0].setVelocity((d.blocks[0].velocity * 0.5).contiguous())
                d.setVelocityResult(torch.zeros_like(d.velocityResult))
                d.UpdateDomainData()
        batch.step()
        for s in sims_ref:
            assert s._PISO_split_step(1)
    for a, b in zip(_state(ref), _state(doms)):
        for x, y in zip(a, b):
            assert torch.equal(x, y)


def test_batch_c4_full_size():
    """The 64-cavity workload: two batched steps against two plain steps of every cavity."""
    doms, ns, visc, dt = configs.cavity_batch_c4(64, 0, torch.float32, "cuda")
    ref, _, _, _ = configs.cavity_batch_c4(64, 0, torch.float32, "cuda")
    sims = [_sim(d, torch.float32, dt) for d in doms]
    batch = SimulationBatch(sims)
    for _ in range(2):
        arr = batch.step()
        assert arr["converged"].all() and arr["is_finite_residual"].all()
    info = batch.describe()
    print(" c4:", info, "cells", batch.total_cells, flush=True)
    assert info["count"] == 64 and info["graph"] == 1
    assert info["pressure_jobs"] >= info["pressure_grid"] > 0
    for d in ref:
        s = _sim(d, torch.float32, dt)
        for _ in range(2):
            assert s._PISO_split_step(1)
    worst = 0.0
    for a, b in zip(_state(ref), _state(doms)):
        for x, y in zip(a, b):
            worst = max(worst, float((x - y).abs().max()) / max(float(x.abs().max()), 1e-30))
    print(" batched vs plain: %.2e" % worst, flush=True)
    assert worst < 1e-4
