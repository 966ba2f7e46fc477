"""Advective outflow (PISOtorch_simulation.update_advective_boundaries as the PRE hook, the way the reference's
tests/test_setups.py:336-343 drives it): a two-block channel with an inflow and an outflow boundary, on a uniform
and on a stretched grid (transformed outflow boundary), stepped by Simulation on the GPU (HIP stages) and by the
FP64 oracle on a host twin that gets the same boundary update before every step.  Block velocities and the
outflow boundary velocity agree to 1e-8 (FP64) / 1e-4 (FP32) relative after five steps, and the domain's boundary
flux stays balanced."""
import pytest
import torch

import scenarios
from compare import rel_err
from oracle import oracle as O
from pict_amd import PISOtorch_simulation as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("coords", [False, True])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_outflow_channel_steps(dtype, coords):
    print("case outflow channel", dtype, "coords" if coords else "uniform", flush=True)
    g = scenarios.outflow_channel(dtype, "cuda", coords)
    out = g.getBlock(1).getBoundary("+x")
    vel_in = torch.tensor([[1.0, 0.0]], dtype=dtype, device="cuda")
    S.balance_boundary_fluxes(g, [out])
    c = O.cpu_twin(g)
    cout = c.getBlock(1).getBoundary("+x")
    dt, steps = 0.05, 5
    tol = 1e-12 if dtype == torch.float64 else 1e-7
    sim = S.Simulation(g, time_step=dt, corrector_steps=2, non_orthogonal=False, advection_tol=tol,
                       pressure_tol=tol if dtype == torch.float64 else 1e-6, advect_passive_scalar=False,
                       prep_fn={"PRE": lambda domain, time_step, **kw: S.update_advective_boundaries(
                           domain, [out], vel_in, time_step.cuda())})
    assert sim.run(steps)
    vin64 = vel_in.cpu().double()
    for _ in range(steps):
        S.update_advective_boundaries(c, [cout], vin64, torch.tensor([dt], dtype=torch.float64))
        O.piso_step(c, dt, corrector_steps=2, advection_tol=1e-12, pressure_tol=1e-12)
    ev = max(rel_err(gb.velocity, cb.velocity) for gb, cb in zip(g.blocks, c.blocks))
    eb = rel_err(out.velocity, cout.velocity)
    flux = float(g.GetBoundaryFluxBalance())
    print(" velocity err %.3e, outflow boundary err %.3e, flux balance %.3e" % (ev, eb, flux), flush=True)
    bound = 1e-8 if dtype == torch.float64 else 1e-4
    assert ev <= bound and eb <= bound
    assert abs(flux) <= (1e-12 if dtype == torch.float64 else 1e-5)
    # the outflow boundary moved away from its initial uniform value
    assert float((out.velocity[0, 0] - 1).abs().max()) > 1e-3
