"""Passive scalars on the GPU against the FP64 oracle (whose scalar stages tests/test_oracle_scalar.py pins by a
Fourier-mode closed form, wall profiles, equality with a velocity component, exactness for linear fields and finite
differences).

Scenarios: the orthogonal ones (cavity, 2-block periodic, rotated connection, 3-D box, inflow/outflow channel) and
the sheared ones (flags 25 and 26), each with 2 channels: random block scalars, a per-channel, static or no scalar
viscosity, per-cell and static boundary scalars and channel 1 Neumann on every second wall.  Tolerances: 1e-11
relative in FP64, 1e-5 in FP32."""
import pytest
import torch

import scenarios
from compare import assert_close, tol_for
from grad_util import OUTPUTS, create_grads, input_tensors
from oracle import oracle as O
from pict_amd import PISOtorch as P

pytestmark = pytest.mark.gpu

DT = 0.05
BUILDERS = dict(scenarios.ALL, **scenarios.NON_ORTHO, outflow_channel=scenarios.outflow_channel)
CASES = [("cavity", 0, "per_channel"), ("periodic_two_block", 0, None), ("rotated_connection", 0, "static"),
         ("box3d", 0, "per_channel"), ("outflow_channel", 0, "per_channel"), ("sheared_walls", 25, "per_channel"),
         ("sheared_two_block_rotated", 26, None), ("sheared_box3d", 25, "static"), ("sheared_periodic_x", 26,
                                                                                   "per_channel")]


def _pair(name, dtype, visc):
    g = scenarios.with_scalars(BUILDERS[name], viscosity=visc, dtype=dtype, device="cuda")
    c = O.cpu_twin(g)
    gen = torch.Generator().manual_seed(23)
    sr = (torch.rand(g.scalarResult.shape, generator=gen, dtype=torch.float64) * 2 - 1) * 0.5
    g.setScalarResult(sr.to(dtype=dtype, device="cuda"))
    c.setScalarResult(sr.to(dtype=dtype).double())
    g.UpdateDomainData()
    return g, c


@pytest.mark.parametrize("name,flags,visc", CASES)
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_scalar_stage_parity(name, flags, visc, dtype):
    print("case", name, flags, visc, dtype, flush=True)
    g, c = _pair(name, dtype, visc)
    tol = tol_for(dtype)
    tdt = torch.tensor([DT], dtype=dtype)
    for ch in range(2):
        P.SetupAdvectionMatrix(g, tdt, flags, forPassiveScalar=True, passiveScalarChannel=ch)
        O.SetupAdvectionMatrix(c, DT, flags, True, ch)
        assert_close(g.C.value, c.C.value, tol, "C values, channel %d" % ch)
        assert_close(g.A, c.A, tol, "A, channel %d" % ch)
    P.SetupAdvectionScalar(g, tdt, flags)
    O.SetupAdvectionScalar(c, DT, flags)
    assert_close(g.scalarRHS, c.scalarRHS, tol, "scalarRHS")
    # copies between the blocks and the global vector
    P.CopyScalarResultToBlocks(g)
    O.CopyScalarResultToBlocks(c)
    for gb, cb in zip(g.blocks, c.blocks):
        assert torch.equal(gb.passiveScalar.cpu().double(), cb.passiveScalar.to(dtype).double())
    g.setScalarResult(torch.zeros_like(g.scalarResult))
    g.UpdateDomainData()
    P.CopyScalarResultFromBlocks(g)
    torch.cuda.synchronize()
    assert torch.equal(g.scalarResult.cpu().double(), c.scalarResult.to(dtype).double())


STAGES = {
    "matrix_ch0": (lambda d, dt, f: P.SetupAdvectionMatrixGrad(d, dt, f, True, 0),
                   lambda d, dt, f: O.SetupAdvectionMatrixGrad(d, dt, f, True, 0),
                   lambda d, dt, f: P.SetupAdvectionMatrix(d, dt, f, True, 0),
                   ["VELOCITY", "BOUNDARY_VELOCITY", "VISCOSITY", "PASSIVE_SCALAR_VISCOSITY"], ["A", "C"]),
    "matrix_ch1": (lambda d, dt, f: P.SetupAdvectionMatrixGrad(d, dt, f, True, 1),
                   lambda d, dt, f: O.SetupAdvectionMatrixGrad(d, dt, f, True, 1),
                   lambda d, dt, f: P.SetupAdvectionMatrix(d, dt, f, True, 1),
                   ["VELOCITY", "BOUNDARY_VELOCITY", "VISCOSITY", "PASSIVE_SCALAR_VISCOSITY"], ["A", "C"]),
    "rhs": (lambda d, dt, f: P.SetupAdvectionScalarGrad(d, dt, f),
            lambda d, dt, f: O.SetupAdvectionScalarGrad(d, dt, f),
            lambda d, dt, f: P.SetupAdvectionScalar(d, dt, f),
            ["PASSIVE_SCALAR", "BOUNDARY_PASSIVE_SCALAR", "BOUNDARY_VELOCITY", "VISCOSITY", "PASSIVE_SCALAR_VISCOSITY",
             "PASSIVE_SCALAR_RESULT"], ["PASSIVE_SCALAR_RHS"]),
}


@pytest.mark.parametrize("stage", list(STAGES))
@pytest.mark.parametrize("name,flags,visc", CASES)
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_scalar_stage_grad_parity(stage, name, flags, visc, dtype):
    print("case", stage, name, flags, visc, dtype, flush=True)
    hip_bwd, ora_bwd, hip_fwd, inputs, outputs = STAGES[stage]
    g, c = _pair(name, dtype, visc)
    tdt = torch.tensor([DT], dtype=dtype)
    hip_fwd(g, tdt, flags)
    gen = torch.Generator().manual_seed(17)
    for d in (g, c):
        create_grads(d)
        d.CreateViscosityGrad()
    for o in outputs:
        ybar = torch.rand(OUTPUTS[o][0](c).shape, generator=gen, dtype=torch.float64) * 2 - 1
        OUTPUTS[o][1](g, ybar.to(dtype=dtype, device="cuda"))
        OUTPUTS[o][1](c, ybar.clone())
    g.UpdateDomainData()
    c._set_changed(False)
    hip_bwd(g, tdt, flags)
    ora_bwd(c, DT, flags)
    torch.cuda.synchronize()
    tol = tol_for(dtype)
    n = 0
    for inp in inputs:
        for (_, gg), (_, gc) in zip(input_tensors(g, inp), input_tensors(c, inp)):
            a, b = gg(), gc()
            assert a is not None and b is not None, inp
            assert a.shape == b.shape, (inp, a.shape, b.shape)
            if float(b.abs().max()) == 0.0:
                assert float(a.abs().max()) == 0.0, inp
                continue
            assert_close(a, b, tol if dtype == torch.float32 else 10 * tol, "%s grad of %s" % (stage, inp))
            n += 1
    assert n > 0


def test_copy_grad_parity():
    g, c = _pair("periodic_two_block", torch.float64, None)
    gen = torch.Generator().manual_seed(5)
    for d in (g, c):
        create_grads(d)
    for gb, cb in zip(g.blocks, c.blocks):
        t = torch.rand(cb.passiveScalar.shape, generator=gen, dtype=torch.float64)
        gb.setPassiveScalarGrad(t.cuda())
        cb.setPassiveScalarGrad(t.clone())
    g.UpdateDomainData()
    P.CopyScalarResultGradFromBlocks(g)
    O.CopyScalarResultGradFromBlocks(c)
    torch.cuda.synchronize()
    assert torch.equal(g.scalarResultGrad.cpu(), c.scalarResultGrad)
    t = torch.rand(c.scalarResultGrad.shape, generator=gen, dtype=torch.float64)
    g.setScalarResultGrad(t.cuda())
    c.setScalarResultGrad(t.clone())
    g.UpdateDomainData()
    P.CopyScalarResultGradToBlocks(g)
    O.CopyScalarResultGradToBlocks(c)
    torch.cuda.synchronize()
    for gb, cb in zip(g.blocks, c.blocks):
        assert torch.equal(gb.passiveScalarGrad.cpu(), cb.passiveScalarGrad)


STEP_CASES = [("cavity", False, "per_channel", torch.float64), ("outflow_channel", False, None, torch.float64),
              ("box3d", False, "static", torch.float64), ("cavity", False, "per_channel", torch.float32),
              ("sheared_two_block_rotated", True, None, torch.float64)]


@pytest.mark.parametrize("name,non_ortho,visc,dtype", STEP_CASES)
def test_driver_steps_with_scalars(name, non_ortho, visc, dtype):
    """Three steps of the driver with passive-scalar advection (Simulation's default advect_passive_scalar=True;
    the orthogonal branch, or the non-orthogonal one with flags 25) against the FP64 oracle's steps."""
    from pict_amd import PISOtorch_simulation as S
    print("case", name, non_ortho, visc, dtype, flush=True)
    g = scenarios.with_scalars(BUILDERS[name], viscosity=visc, dtype=dtype, device="cuda")
    c = O.cpu_twin(g)
    # the non-orthogonal pressure matrix is not symmetric: its CG (the reference's) converges to 1e-11 here
    tol_solve = (1e-11 if non_ortho else 1e-12) if dtype == torch.float64 else 1e-6
    sim = S.Simulation(g, time_step=DT, corrector_steps=2, non_orthogonal=non_ortho, advection_tol=tol_solve,
                       pressure_tol=tol_solve)
    assert sim.advect_passive_scalar and not sim._native_step_eligible()
    steps = 3
    assert sim.run(steps)
    ora_tol = 1e-11 if non_ortho else 1e-13
    for _ in range(steps):
        O.piso_step(c, DT, corrector_steps=2, advection_tol=ora_tol, pressure_tol=ora_tol, non_orthogonal=non_ortho)
    bound = (1e-6 if non_ortho else 1e-9) if dtype == torch.float64 else 1e-4
    for f in ("passiveScalar", "velocity", "pressure"):
        e = max(float((getattr(gb, f).cpu().double() - getattr(cb, f)).abs().max()) for gb, cb in
                zip(g.blocks, c.blocks)) / max(float(getattr(cb, f).abs().max()) for cb in c.blocks)
        print("  %s %.3e" % (f, e), flush=True)
        assert e <= bound, (f, e)


def test_differentiable_step_scalar_gradient_matches_finite_differences():
    """Gradient of a loss of the advected scalar after two differentiable steps (PISOtorch_diff: scalar matrix,
    scalar RHS, the multi-right-hand-side adjoint solve) w.r.t. the initial block scalars and the boundary scalars,
    against central differences of the forward steps (FP64, solves at 1e-12)."""
    from pict_amd import PISOtorch_simulation as S

    def make():
        d = scenarios.with_scalars(BUILDERS["outflow_channel"], viscosity="static", neumann=False,
                                   dtype=torch.float64, device="cuda")
        return d

    def run(d, differentiable):
        sim = S.Simulation(d, time_step=DT, corrector_steps=2, non_orthogonal=False, advection_tol=1e-12,
                           pressure_tol=1e-12, differentiable=differentiable)
        assert sim._PISO_split_step(2)

    gen = torch.Generator().manual_seed(29)
    d = make()
    s0 = [b.passiveScalar.detach().clone() for b in d.blocks]
    target = [torch.rand(s.shape, generator=gen, dtype=torch.float64).cuda() for s in s0]
    ss = [s.clone().requires_grad_(True) for s in s0]
    for b, s in zip(d.blocks, ss):
        b.setPassiveScalar(s)
    d.UpdateDomainData()
    run(d, True)
    loss = sum(((b.passiveScalar - t) ** 2).sum() for b, t in zip(d.blocks, target))
    loss.backward()
    grads = [s.grad.detach() for s in ss]
    dirs = [torch.rand(s.shape, generator=gen, dtype=torch.float64).cuda() * 2 - 1 for s in s0]
    an = sum(float((gr * v).sum()) for gr, v in zip(grads, dirs))
    eps = 1e-6
    vals = []
    for sgn in (1, -1):
        dd = make()
        for b, s, v in zip(dd.blocks, s0, dirs):
            b.setPassiveScalar((s + sgn * eps * v).contiguous())
        dd.UpdateDomainData()
        run(dd, False)
        vals.append(sum(float(((b.passiveScalar - t) ** 2).sum()) for b, t in zip(dd.blocks, target)))
    fd = (vals[0] - vals[1]) / (2 * eps)
    rel = abs(fd - an) / max(abs(fd), 1e-12)
    print("  adjoint %.10e fd %.10e rel %.2e" % (an, fd, rel), flush=True)
    assert rel < 1e-4


def test_reference_defaults_run():
    """A driver script with the reference's defaults runs: Domain with its default passiveScalarChannels=1, a
    transformed grid, and Simulation(domain, time_step) with its defaults (non_orthogonal=True,
    advect_passive_scalar=True); the scalar stays within its initial and boundary range (a discrete maximum
    principle for this diffusion-dominated cavity) and the velocity stays near divergence-free."""
    from pict_amd import PISOtorch_simulation as S
    from pict_amd import configs
    d = P.Domain(2, torch.tensor([0.05], dtype=torch.float32), "defaults", torch.float32, "cuda")
    assert d.getPassiveScalarChannels() == 1
    b = d.CreateBlock(vertexCoordinates=configs.uniform_coords(24, 24, dtype=torch.float32, device="cuda"))
    b.CloseAllBoundaries()
    b.CloseBoundary("+y", velocity=torch.tensor([[1.0, 0.0]], device="cuda"),
                    passiveScalar=torch.tensor([[1.0]], device="cuda"))
    d.PrepareSolve()
    sim = S.Simulation(d, time_step=0.01)
    assert sim.non_orthogonal and sim.advect_passive_scalar
    assert sim.run(5)
    s = b.passiveScalar
    assert torch.isfinite(s).all() and float(s.min()) >= -1e-5 and float(s.max()) <= 1.0 + 1e-5
    assert float(s.max()) > 0.01  # the scalar entered from the lid
    # max |div u| 1.4e-3 measured: FP32 at the driver's default solver tolerance, largest at the lid's corners
    div = float(P.ComputeVelocityDivergence(d).abs().max())
    assert div < 1e-2, div
