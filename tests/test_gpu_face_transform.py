"""useFaceTransform in the pressure setup (computeFluxesWithFaceTransforms PISO_multiblock_cuda_kernel.cu:1703-1785
in the pressure-RHS divergence :5374-5419; getLaplaceCoefficientOrthogonalFace :1226-1241 in the pressure matrix
:4857-4865) on the GPU against the FP64 oracle (tests/test_oracle_face_transform.py pins it), on every scenario with
vertex coordinates, and the non-orthogonal terms with face coefficients on the sheared scenarios: FP64 1e-11, FP32
1e-5 relative (the divergence, a difference of fluxes, 10x that).  Blocks without face transforms: the matrix takes
the coefficient 1 as in the reference, the divergence is refused."""
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
def test_face_transform_pressure_setup_matches_the_oracle(name, dtype):
    print("case", name, dtype, flush=True)
    g = CASES[name](dtype=dtype, device="cuda")
    c = O.cpu_twin(g)
    tol = tol_for(dtype)
    dt = 0.05
    tdt = torch.tensor([dt], dtype=dtype)
    P.SetupAdvectionMatrix(g, tdt, 0)
    O.SetupAdvectionMatrix(c, dt)
    P.SetupPressureMatrix(g, tdt, 0, True)
    O.SetupPressureMatrix(c, dt, 0, True)
    assert_close(g.P.value, c.P.value, tol, "pressure matrix with face transforms")
    H = _rand(g.pressureRHS.shape, 23, 0.5)
    g.setPressureRHS(H.to(dtype=dtype, device="cuda"))
    c.setPressureRHS(H.to(dtype=dtype).double())
    g.UpdateDomainData()
    c.UpdateDomainData()
    faces = all(b.hasFaceTransform() for b in g.getBlocks())
    for tsn in (False, True):
        if faces:
            P.SetupPressureRHSdiv(g, tdt, 0, True, tsn)
            O.SetupPressureRHSdiv(c, dt, 0, True, tsn)
            assert_close(g.pressureRHSdiv, c.pressureRHSdiv, 10 * tol, "divergence with face transforms")
        else:
            with pytest.raises(P.PictError):
                P.SetupPressureRHSdiv(g, tdt, 0, True, tsn)
            with pytest.raises(RuntimeError):
                O.SetupPressureRHSdiv(c, dt, 0, True, tsn)


@pytest.mark.parametrize("name", list(scenarios.NON_ORTHO))
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_non_orthogonal_face_terms_match_the_oracle(name, dtype):
    """The non-orthogonal pressure terms with face coefficients (:1962-1972, :3128-3146; flags 25) and the VJP of
    the explicit RHS terms with them (:3294-3312, :3364-3368) against the oracle."""
    print("case", name, dtype, flush=True)
    g = scenarios.NON_ORTHO[name](dtype=dtype, device="cuda")
    c = O.cpu_twin(g)
    tol = tol_for(dtype)
    dt = 0.05
    tdt = torch.tensor([dt], dtype=dtype)
    P.SetupAdvectionMatrix(g, tdt, 0)
    O.SetupAdvectionMatrix(c, dt)
    P.SetupPressureMatrix(g, tdt, 25, True)
    O.SetupPressureMatrix(c, dt, 25, True)
    assert_close(g.P.value, c.P.value, tol, "non-orthogonal pressure matrix with face coefficients")
    H = _rand(g.pressureRHS.shape, 28, 0.5)
    pr = _rand(g.pressureRHSdiv.shape, 29)
    g.setPressureRHS(H.to(dtype=dtype, device="cuda"))
    c.setPressureRHS(H.to(dtype=dtype).double())
    g.setPressureResult(pr.to(dtype=dtype, device="cuda"))
    c.setPressureResult(pr.to(dtype=dtype).double())
    g.UpdateDomainData()
    c.UpdateDomainData()
    P.SetupPressureRHSdiv(g, tdt, 25, True)
    O.SetupPressureRHSdiv(c, dt, 25, True)
    assert_close(g.pressureRHSdiv, c.pressureRHSdiv, 10 * tol, "divergence with non-orthogonal face terms")
    # the VJP of the non-orthogonal RHS terms with face coefficients, against the oracle's (pinned by finite
    # differences); the matrix VJP is the cell-metric one, as in the reference
    from grad_util import create_grads, input_tensors
    create_grads(g)
    create_grads(c)
    ybar = _rand(g.pressureRHSdiv.shape, 31)
    g.setPressureRHSdivGrad(ybar.to(dtype=dtype, device="cuda"))
    c.setPressureRHSdivGrad(ybar.to(dtype=dtype).double())
    g.UpdateDomainData()
    c._set_changed(False)
    P.SetupPressureRHSdivGrad(g, tdt, 25, True)
    O.SetupPressureRHSdivGrad(c, dt, 25, True)
    torch.cuda.synchronize()
    for inp in ("PRESSURE_RESULT", "A", "PRESSURE_RHS"):
        for (_, gg), (_, gc) in zip(input_tensors(g, inp), input_tensors(c, inp)):
            assert_close(gg(), gc(), tol if dtype == torch.float32 else 10 * tol, "face VJP of %s" % inp)


@pytest.mark.parametrize("name", ["rotated_connection", "box3d", "sheared_two_block_rotated"])
def test_face_flag_leaves_the_orthogonal_vjps_unchanged(name):
    """The reference's pressure-matrix and divergence VJPs do not use useFaceTransform outside the non-orthogonal
    terms (:4966-5100, :5422-5452): with the flag, the VJP of the pressure correction equals the one without it (to
    rounding: the scatters accumulate with atomics, whose order varies between runs)."""
    from grad_util import OUTPUTS, create_grads, input_tensors
    print("case", name, flush=True)
    tdt = torch.tensor([0.05], dtype=torch.float64)
    grads = []
    for face in (False, True):
        g = CASES[name](dtype=torch.float64, device="cuda")
        P.SetupAdvectionMatrix(g, tdt, 0)
        P.SetupAdvectionVelocity(g, tdt, 0)
        g.setVelocityResult(_rand(g.velocityResult.shape, 27, 0.5).cuda())
        g.UpdateDomainData()
        P.SetupPressureCorrection(g, tdt, 0)
        create_grads(g)
        for k, o in enumerate(("P", "PRESSURE_RHS", "PRESSURE_RHS_DIV")):
            OUTPUTS[o][1](g, _rand(OUTPUTS[o][0](g).shape, 30 + k).cuda())
        g.UpdateDomainData()
        P.SetupPressureCorrectionGrad(g, tdt, 0, face)
        torch.cuda.synchronize()
        grads.append([gg().clone() for inp in ("VELOCITY_RESULT", "A", "C", "VELOCITY")
                      for _, gg in input_tensors(g, inp)])
    assert len(grads[0]) > 0
    for a, b in zip(*grads):
        assert_close(a, b, 1e-13, "VJP with and without useFaceTransform")
