"""CorrectVelocity versions 4, 5 and 6 (PISO_multiblock_cuda_kernel.cu:6052-6160, :6013-6049 with :991-1124; the
version switch :6163-6203) on the GPU against the FP64 oracle (tests/test_oracle_correctors.py pins the oracle), on
every scenario: orthogonal and sheared transforms, periodic faces, connections with permuted and inverted axes,
walls, 2-D and 3-D, with and without time-step normalisation; other version numbers take the finite-difference
corrector.  Tolerances as for the stages: FP64 1e-11, FP32 1e-5 relative (max|a-b| / max|b|)."""
import pytest
import torch

import scenarios
from compare import assert_close, tol_for
from oracle import oracle as O
from pict_amd import PISOtorch as P

pytestmark = pytest.mark.gpu

CASES = dict(scenarios.ALL, **scenarios.NON_ORTHO, outflow_channel=scenarios.outflow_channel)


def _rand(shape, seed, scale=1.0):
    g = torch.Generator().manual_seed(seed)
    return (torch.rand(shape, generator=g, dtype=torch.float64) * 2 - 1) * scale


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_correctors_match_the_oracle(name, dtype):
    print("case", name, dtype, flush=True)
    g = CASES[name](dtype=dtype, device="cuda")
    c = O.cpu_twin(g)
    tol = tol_for(dtype)
    dt = 0.05
    tdt = torch.tensor([dt], dtype=dtype)
    P.SetupAdvectionMatrix(g, tdt, 0)
    O.SetupAdvectionMatrix(c, dt)
    H = _rand(g.pressureRHS.shape, 21, 0.5)
    g.setPressureRHS(H.to(dtype=dtype, device="cuda"))
    c.setPressureRHS(H.to(dtype=dtype).double())
    for i, (gb, cb) in enumerate(zip(g.blocks, c.blocks)):
        p = _rand(gb.pressure.shape, 30 + i)
        gb.setPressure(p.to(dtype=dtype, device="cuda"))
        cb.setPressure(p.to(dtype=dtype).double())
    g.UpdateDomainData()
    c.UpdateDomainData()
    for version, tsn in ((4, False), (5, False), (5, True), (6, False), (6, True), (3, True)):
        P.CorrectVelocity(g, tdt, version, tsn)
        O.CorrectVelocity(c, dt, version, tsn)
        assert_close(g.velocityResult, c.velocityResult, tol, "version %d (time step norm %s)" % (version, tsn))


def test_face_metric_corrector_needs_face_transforms():
    """A transformed block without face transforms has no boundary transforms on its walls: versions 5 and 6 refuse
    it instead of reading them."""
    g = scenarios.cavity(dtype=torch.float32, device="cuda")
    b = g.getBlocks()[0]
    b.setTransform(b.transform.clone())
    b.CloseAllBoundaries()
    g.PrepareSolve()
    tdt = torch.tensor([0.05], dtype=torch.float32)
    P.SetupAdvectionMatrix(g, tdt, 0)
    for version in (5, 6):
        with pytest.raises(P.PictError):
            P.CorrectVelocity(g, tdt, version)
    P.CorrectVelocity(g, tdt, 4)
    P.CorrectVelocity(g, tdt, 1)


@pytest.mark.parametrize("corrector,version", [("FVM_CENTER", 5), ("FVM_FACE", 6)])
@pytest.mark.parametrize("name,non_ortho", [("rotated_connection", False), ("sheared_two_block_rotated", True)])
def test_simulation_steps_with_finite_volume_correctors(corrector, version, name, non_ortho):
    """Simulation(velocity_corrector=...) on the GPU against the FP64 oracle's step with the same corrector
    (PISOtorch_simulation.py:450, the corrector call :1305): three FP64 steps to 1e-8 relative."""
    from pict_amd import PISOtorch_simulation as S
    print("case", name, corrector, flush=True)
    dt = 0.02
    g = CASES[name](dtype=torch.float64, device="cuda")
    c = O.cpu_twin(g)
    sim = S.Simulation(g, time_step=dt, corrector_steps=2, advection_tol=1e-12, pressure_tol=1e-12,
                       advect_passive_scalar=False, non_orthogonal=non_ortho, velocity_corrector=corrector)
    assert sim.run(3)
    for _ in range(3):
        O.piso_step(c, dt, corrector_steps=2, advection_tol=1e-12, pressure_tol=1e-12, non_orthogonal=non_ortho,
                    corrector_version=version)
    for gb, cb in zip(g.blocks, c.blocks):
        assert_close(gb.velocity, cb.velocity, 1e-8, "velocity after 3 steps")
    assert_close(g.pressureResult, c.pressureResult, 1e-8, "pressure after 3 steps")
