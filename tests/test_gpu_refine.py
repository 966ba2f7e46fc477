"""FP64 iterative refinement of FP32 solves (PISOtorch.SolveLinearRefined, Simulation(pressure_refine_tol=...)):
the FP64 residual of the FP32-valued system reaches the requested 1e-10 relative (checked in FP64 on the host with
scipy), the solution agrees with a direct FP64 solve of the same values, and a refined FP32 PISO step is closer to
the FP64 step than the plain FP32 step."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla
import torch

import scenarios
from pict_amd import PISOtorch as P
from pict_amd import PISOtorch_simulation as S
from test_gpu_solver import periodic_laplacian, rhs, to_pict

pytestmark = pytest.mark.gpu


def _rel_residual(m, b, x):
    r = b.astype(np.float64) - m @ x
    return np.linalg.norm(r) / np.linalg.norm(b.astype(np.float64))


@pytest.mark.parametrize("grid", [(256, 256), (48, 32, 32)])
def test_refined_solve_reaches_fp64_residual(grid):
    m = periodic_laplacian(grid, 0.01)
    m.data = m.data.astype(np.float32).astype(np.float64)  # the FP32 values, exactly
    A = to_pict(m, torch.float32)
    b = rhs(m.shape[0], 5, torch.float32)
    x = torch.zeros(m.shape[0], dtype=torch.float64, device="cuda")
    info = P.SolveLinearRefined(A, b, x, 5000, 1e-5, 1e-10, 8)
    bh = b.cpu().numpy().astype(np.float64)
    res = _rel_residual(m, bh, x.cpu().numpy())
    print(" refined: sweeps-level residual %.3e (reported %.3e), inner its %d" % (res, info.finalResidual,
                                                                                info.usedIterations), flush=True)
    assert info.converged and res < 1e-10
    xd = spla.spsolve(m.tocsc(), bh)
    assert np.abs(x.cpu().numpy() - xd).max() / np.abs(xd).max() < 1e-8
    # the plain FP32 solve stops far above that
    x32 = torch.zeros_like(b)
    P.SolveLinear(A, b, x32, 5000, 1e-7, P.ConvergenceCriterion.NORM2_NORMALIZED, False, False, 0, False, False,
                  False, False)
    assert _rel_residual(m, bh, x32.double().cpu().numpy()) > 1e-8


def test_refined_solve_singular_pressure_system():
    m = periodic_laplacian((128, 96), 0.0)
    A = to_pict(m, torch.float32)
    # a right-hand side whose FP32 values sum to exactly zero (multiples of 2^-20): the singular system is
    # consistent in the matrix dtype, not only in FP64
    rng = np.random.default_rng(8)
    half = rng.integers(-(1 << 20), 1 << 20, m.shape[0] // 2).astype(np.int64)
    k = rng.permutation(np.concatenate([half, -half]))
    bh = k.astype(np.float64) * 2.0 ** -20
    b32 = torch.tensor(bh, dtype=torch.float32, device="cuda")
    assert float(b32.double().sum()) == 0.0
    x = torch.zeros(m.shape[0], dtype=torch.float64, device="cuda")
    info = P.SolveLinearRefined(A, b32, x, 5000, 1e-5, 1e-10, 8)
    res = _rel_residual(m, bh, x.cpu().numpy())
    print(" singular: residual %.3e" % res, flush=True)
    assert info.converged and res < 1e-10


def _step_pressure(name, dtype, refine, inputs=None):
    d = scenarios.ALL[name](dtype=dtype, device="cuda")
    if inputs is not None:
        for b, v in zip(d.blocks, inputs):
            # a copy: the step updates the block velocity in place, and .to() of the same dtype aliases the input
            b.setVelocity(v.to(dtype).clone().contiguous())
        d.UpdateDomainData()
    sim = S.Simulation(d, time_step=0.05, corrector_steps=2, non_orthogonal=False,
                       advection_tol=1e-12 if dtype == torch.float64 else 1e-6,
                       pressure_tol=1e-12 if dtype == torch.float64 else 1e-6, advect_passive_scalar=False,
                       pressure_refine_tol=1e-10 if refine else None)
    assert sim._PISO_split_step(1)
    return [b.pressure.detach().double().clone() for b in d.blocks], [b.velocity.detach().clone() for b in d.blocks]


@pytest.mark.parametrize("name", ["periodic_two_block", "cavity"])
def test_refined_fp32_step_is_closer_to_fp64(name):
    d0 = scenarios.ALL[name](dtype=torch.float32, device="cuda")
    v0 = [b.velocity.detach().clone() for b in d0.blocks]
    p64, _ = _step_pressure(name, torch.float64, False, v0)
    p32, _ = _step_pressure(name, torch.float32, False, v0)
    p32r, _ = _step_pressure(name, torch.float32, True, v0)

    def err(p):
        return max(float((a - b).abs().max()) for a, b in zip(p, p64)) / max(float(b.abs().max()) for b in p64)

    e_plain, e_ref = err(p32), err(p32r)
    print(" pressure vs FP64 step: plain %.3e refined %.3e" % (e_plain, e_ref), flush=True)
    assert e_ref <= e_plain and e_ref < 1e-4
