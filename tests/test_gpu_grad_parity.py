"""HIP backward stages vs the FP64 oracle's (pinned by finite differences in test_oracle_grad.py) on identical
inputs: the forward state of the GPU domain (C, A, velocity result, pressure RHS) is copied into the oracle twin
before both run the same stage VJP on the same output cotangents.  Tolerances as for the forward stages: 1e-11
relative in FP64, 1e-5 in FP32 (max|a-b| / max|b| per gradient tensor)."""
import pytest
import torch

import scenarios
from compare import assert_close, rel_err, tol_for
from grad_util import OUTPUTS, create_grads, input_tensors
from oracle import oracle as O
from pict_amd import PISOtorch as P

pytestmark = pytest.mark.gpu

DT = 0.05
GPU = {
    "SetupAdvectionMatrix": (lambda d, dt: P.SetupAdvectionMatrixGrad(d, dt, 0), O.SetupAdvectionMatrixGrad,
                             ["VELOCITY", "BOUNDARY_VELOCITY", "VISCOSITY", "VISCOSITY_BLOCK"], ["A", "C"]),
    "SetupAdvectionVelocity": (lambda d, dt: P.SetupAdvectionVelocityGrad(d, dt, 0), O.SetupAdvectionVelocityGrad,
                               ["VELOCITY", "VELOCITY_SOURCE", "BOUNDARY_VELOCITY", "VISCOSITY", "VISCOSITY_BLOCK"],
                               ["VELOCITY_RHS"]),
    "SetupPressureMatrix": (P.SetupPressureMatrixGrad, O.SetupPressureMatrixGrad, ["A"], ["P"]),
    "SetupPressureRHSdiv": (P.SetupPressureRHSdivGrad, O.SetupPressureRHSdivGrad,
                            ["PRESSURE_RHS", "BOUNDARY_VELOCITY"], ["PRESSURE_RHS_DIV"]),
    "SetupPressureRHS": (P.SetupPressureRHSGrad, O.SetupPressureRHSGrad,
                         ["VELOCITY", "VELOCITY_SOURCE", "VELOCITY_RESULT", "BOUNDARY_VELOCITY", "VISCOSITY",
                          "VISCOSITY_BLOCK", "C", "A"], ["PRESSURE_RHS", "PRESSURE_RHS_DIV"]),
    "SetupPressureCorrection": (P.SetupPressureCorrectionGrad, O.SetupPressureCorrectionGrad,
                                ["VELOCITY", "VELOCITY_SOURCE", "VELOCITY_RESULT", "BOUNDARY_VELOCITY", "VISCOSITY",
                                 "VISCOSITY_BLOCK", "C", "A"], ["P", "PRESSURE_RHS", "PRESSURE_RHS_DIV"]),
    "CorrectVelocity": (lambda d, dt: P.CorrectVelocityGrad(d, dt, False), O.CorrectVelocityGrad,
                        ["A", "PRESSURE", "PRESSURE_RHS"], ["VELOCITY_RESULT"]),
}


@pytest.mark.parametrize("stage", list(GPU))
@pytest.mark.parametrize("name", list(scenarios.ALL))
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_stage_grad_parity(stage, name, dtype):
    print("case", stage, name, dtype, flush=True)
    hip_bwd, ora_bwd, inputs, outputs = GPU[stage]
    g = scenarios.ALL[name](dtype=dtype, device="cuda")
    c = O.cpu_twin(g)
    tdt = torch.tensor([DT], dtype=dtype)
    gen = torch.Generator().manual_seed(11)
    P.SetupAdvectionMatrix(g, tdt, 0)
    P.SetupAdvectionVelocity(g, tdt, 0)
    g.setVelocityResult(((torch.rand(g.velocityResult.shape, generator=gen, dtype=torch.float64) * 2 - 1) * 0.5)
                        .to(dtype=dtype, device="cuda"))
    g.UpdateDomainData()
    P.SetupPressureCorrection(g, tdt, 0)
    # identical forward state on the oracle side
    for name_, src, dst in (("C", g.C.value, c.C.value), ("A", g.A, c.A), ("vr", g.velocityResult, c.velocityResult),
                            ("prhs", g.pressureRHS, c.pressureRHS), ("P", g.P.value, c.P.value)):
        dst.copy_(src.detach().cpu().double())
    ybar = {o: ((torch.rand(OUTPUTS[o][0](c).shape, generator=gen, dtype=torch.float64) * 2 - 1)) for o in outputs}
    create_grads(g)
    create_grads(c)
    for o in outputs:
        OUTPUTS[o][1](g, ybar[o].to(dtype=dtype, device="cuda"))
        OUTPUTS[o][1](c, ybar[o].clone())
    g.UpdateDomainData()
    c._set_changed(False)
    hip_bwd(g, tdt)
    ora_bwd(c, DT)
    torch.cuda.synchronize()
    tol = tol_for(dtype)
    n = 0
    for inp in inputs:
        for (tg, gg), (tc, gc) in zip(input_tensors(g, inp), input_tensors(c, inp)):
            a, b = gg(), gc()
            assert a.shape == b.shape, (inp, a.shape, b.shape)
            print("  %s %.3e" % (inp, rel_err(a, b)), flush=True)
            assert_close(a, b, tol if dtype == torch.float32 else 10 * tol, "%s grad of %s" % (stage, inp))
            n += 1
    assert n > 0
