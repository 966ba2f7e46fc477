"""HIP path vs the FP64 oracle on the same seeded inputs, stage by stage and over whole PISO steps.

Tolerances: FP64 kernels agree with the oracle to 1e-11 relative (rounding level);
FP32 kernels to 1e-5 relative, measured as max|a-b|/max|b|.
"""
import pytest
import torch

import scenarios
from compare import assert_close, rel_err, tol_for
from oracle import oracle as O
from pict_amd import PISOtorch as P

pytestmark = pytest.mark.gpu

DTYPES = [torch.float64, torch.float32]


def _rand_like(t, seed, scale=1.0):
    g = torch.Generator().manual_seed(seed)
    return ((torch.rand(t.shape, generator=g, dtype=torch.float64) * 2 - 1) * scale)


@pytest.mark.parametrize("name", list(scenarios.ALL))
@pytest.mark.parametrize("dtype", DTYPES)
def test_stage_parity(name, dtype):
    print("case", name, dtype, flush=True)
    g = scenarios.ALL[name](dtype=dtype, device="cuda")
    c = O.cpu_twin(g)
    tol = tol_for(dtype)
    # exact CSR pattern and indexing
    assert torch.equal(g.C.row.cpu(), c.C.row), "CSR row pointers differ"
    assert torch.equal(g.C.index.cpu(), c.C.index), "CSR column indices differ"
    dt = 0.05
    tdt = torch.tensor([dt], dtype=dtype)
    P.SetupAdvectionMatrix(g, tdt, 0)
    O.SetupAdvectionMatrix(c, dt)
    assert_close(g.C.value, c.C.value, tol, "C values")
    assert_close(g.A, c.A, tol, "A")
    P.SetupAdvectionVelocity(g, tdt, 0)
    O.SetupAdvectionVelocity(c, dt)
    assert_close(g.velocityRHS, c.velocityRHS, tol, "velocityRHS")
    P.SetupAdvectionVelocity(g, tdt, 0, True)
    O.SetupAdvectionVelocity(c, dt, 0, True)
    assert_close(g.velocityRHS, c.velocityRHS, tol, "velocityRHS with pressure gradient")
    vr = _rand_like(g.velocityResult, 11, 0.5)
    g.setVelocityResult(vr.to(dtype=dtype, device="cuda"))
    c.setVelocityResult(vr.to(dtype=dtype).double())
    g.UpdateDomainData()
    P.SetupPressureCorrection(g, tdt, 0)
    O.SetupPressureCorrection(c, dt)
    assert_close(g.P.value, c.P.value, tol, "P values")
    assert_close(g.pressureRHS, c.pressureRHS, tol, "pressureRHS")
    assert_close(g.pressureRHSdiv, c.pressureRHSdiv, 10 * tol, "pressureRHSdiv")
    for gb, cb in zip(g.blocks, c.blocks):
        p = _rand_like(gb.pressure, 5, 1.0)
        gb.setPressure(p.to(dtype=dtype, device="cuda"))
        cb.setPressure(p.to(dtype=dtype).double())
    g.UpdateDomainData()
    P.CorrectVelocity(g, tdt, 1)
    O.CorrectVelocity(c, dt)
    assert_close(g.velocityResult, c.velocityResult, tol, "corrected velocity")
    dg = P.ComputeVelocityDivergence(g)
    dc = O.ComputeVelocityDivergence(c)
    assert_close(dg, dc, 10 * tol, "velocity divergence")


@pytest.mark.parametrize("name", ["cavity", "periodic_two_block", "rotated_connection", "box3d"])
@pytest.mark.parametrize("variant", ["bicg", "cg", "cg_textbook"])
def test_solver_parity(name, variant, monkeypatch):
    """Same matrix, same RHS: HIP solver vs oracle solver (FP64).  The textbook kernels follow the oracle's
    iteration counts within 1; the single-reduction CG (same iterates in exact arithmetic) within 3%+2."""
    print("case", name, variant, flush=True)
    use_bicg = variant == "bicg"
    monkeypatch.setenv("PICT_CG_TEXTBOOK", "1" if variant == "cg_textbook" else "0")
    g = scenarios.ALL[name](dtype=torch.float64, device="cuda")
    dt = torch.tensor([0.05], dtype=torch.float64)
    P.SetupAdvectionMatrix(g, dt, 0)
    P.SetupAdvectionVelocity(g, dt, 0)
    if use_bicg:
        A, b = g.C, g.velocityRHS
    else:
        P.SetupPressureCorrection(g, dt, 0)
        A, b = g.P, g.pressureRHSdiv
        b = b - b.mean()  # compatible RHS for the singular Poisson matrix
    Acpu = P.CSRmatrix(A.value.cpu(), A.index.cpu(), A.row.cpu())
    for transpose in (False, True):
        xg = torch.zeros_like(b)
        ig = P.SolveLinear(A, b.contiguous(), xg, 2000, 1e-11, P.ConvergenceCriterion.NORM2_NORMALIZED, use_bicg,
                           False, 0, transpose, False, False, False)
        xc = torch.zeros(b.shape, dtype=torch.float64)
        ic = O.SolveLinear(Acpu, b.cpu().contiguous(), xc, 2000, 1e-11, 1, use_bicg, 0, transpose)
        print(name, variant, "T" if transpose else "", [i.usedIterations for i in ig],
              [i.usedIterations for i in ic], rel_err(xg, xc), flush=True)
        for a_, c_ in zip(ig, ic):
            assert a_.converged and c_.converged
            slack = 1 if variant != "cg" else 2 + 0.03 * c_.usedIterations
            assert abs(a_.usedIterations - c_.usedIterations) <= slack
        assert_close(xg, xc, 1e-8, "solution")


@pytest.mark.parametrize("name", list(scenarios.ALL))
@pytest.mark.parametrize("dtype", DTYPES)
def test_piso_steps_parity(name, dtype):
    """Three full PISO steps (orthogonal driver order): HIP vs FP64 oracle."""
    print("case", name, dtype, flush=True)
    from pict_amd import PISOtorch_simulation as S
    g = scenarios.ALL[name](dtype=dtype, device="cuda")
    c = O.cpu_twin(g)
    tol_solve = 1e-12 if dtype == torch.float64 else 1e-7
    sim = S.Simulation(g, time_step=0.05, corrector_steps=2, non_orthogonal=False, advection_tol=tol_solve,
                       pressure_tol=tol_solve, advect_passive_scalar=False)
    steps = 3
    sim.run(steps)
    for _ in range(steps):
        O.piso_step(c, 0.05, corrector_steps=2, advection_tol=1e-12, pressure_tol=1e-12)
    ev = max(rel_err(gb.velocity, cb.velocity) for gb, cb in zip(g.blocks, c.blocks))
    ep = rel_err(g.pressureResult, c.pressureResult)
    print(name, dtype, "velocity err %.3e pressure err %.3e" % (ev, ep), flush=True)
    bound = 1e-8 if dtype == torch.float64 else 1e-4
    assert ev <= bound and ep <= bound


def test_transform_parity():
    import scenarios as S_
    for dtype in DTYPES:
        coords = S_.stretched_coords(7, 5, dtype, "cuda", 1.2)
        coords = coords + 0.03 * torch.sin(7 * coords.flip(1))  # non-orthogonal cells
        tg = P.CoordsToTransforms(coords)
        tc = O.CoordsToTransforms(coords.cpu().double())
        assert_close(tg, tc, tol_for(dtype), "cell transforms")
        fg = P.CoordsToFaceTransforms(coords)
        fc = O.CoordsToFaceTransforms(coords.cpu().double())
        assert_close(fg, fc, tol_for(dtype), "face transforms")


@pytest.mark.parametrize("name", ["cavity", "periodic_two_block_coords", "rotated_connection", "box3d"])
@pytest.mark.parametrize("dtype", DTYPES)
def test_native_step_matches_python_driver(name, dtype):
    """pict_piso_step (the orthogonal step as one library call) is bitwise the stage-by-stage Python driver."""
    print("case", name, dtype, flush=True)
    from pict_amd import PISOtorch_simulation as S
    out = []
    for native in (True, False):
        g = scenarios.ALL[name](dtype=dtype, device="cuda")
        sim = S.Simulation(g, time_step=0.05, corrector_steps=2, non_orthogonal=False, advection_tol=1e-10 if
                           dtype == torch.float64 else 1e-6, pressure_tol=1e-10 if dtype == torch.float64 else 1e-6,
                           advect_passive_scalar=False, native_step=native)
        assert sim._native_step_eligible() == native
        sim.record_solver_infos = True
        assert sim._PISO_split_step(2)
        out.append(([b.velocity.clone() for b in g.blocks], [b.pressure.clone() for b in g.blocks],
                    [(k, [i.usedIterations for i in inf]) for k, inf in sim.solver_infos]))
    (va, pa, ia), (vb, pb, ib) = out
    assert ia == ib, (ia, ib)
    for a_, b_ in zip(va + pa, vb + pb):
        assert torch.equal(a_, b_)
