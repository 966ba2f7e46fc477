"""ComputeSpatialVelocityGradients and SGSviscosityIncompressibleSmagorinsky (PISOtorch.cpp:530, :540-541) on the
GPU against the FP64 oracle (tests/test_oracle_sgs.py pins the oracle), on every scenario: orthogonal and
non-orthogonal transforms, periodic faces, connections with permuted axes, fixed walls, 2-D and 3-D.
Tolerances as for the stages: FP64 1e-11, FP32 1e-5 relative (max|a-b| / max|b|).  The per-block tensors the
API returns are compared with the oracle's global vectors cut at each block's offset."""
import pytest
import torch

import scenarios
from compare import assert_close, tol_for
from oracle import oracle as O
from pict_amd import PISOtorch as P

pytestmark = pytest.mark.gpu

CASES = dict(scenarios.ALL, **scenarios.NON_ORTHO, outflow_channel=scenarios.outflow_channel)


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_gradients_and_sgs_match_the_oracle(name, dtype):
    print("case", name, dtype, flush=True)
    g = CASES[name](dtype=dtype, device="cuda")
    c = O.cpu_twin(g)
    tol = tol_for(dtype)
    D = g.getSpatialDims()
    grads = P.ComputeSpatialVelocityGradients(g)
    ref = O.SpatialVelocityGradients(c)  # [c, i, N]
    assert len(grads) == g.getNumBlocks()
    for bi, b in enumerate(g.getBlocks()):
        assert len(grads[bi]) == D
        n = b.pressure.numel()
        for comp in range(D):
            t = grads[bi][comp]
            assert t.shape == b.velocity.shape and t.dtype == dtype and t.is_cuda
            want = ref[comp, :, b.globalOffset:b.globalOffset + n].reshape(b.velocity.shape)
            assert_close(t, want, tol, "block %d gradient of u_%d" % (bi, comp))
    coef = torch.tensor([0.17], dtype=dtype)
    nu = P.SGSviscosityIncompressibleSmagorinsky(g, coef)
    ref_nu = O.SGSViscositySmagorinsky(c, 0.17)
    for bi, b in enumerate(g.getBlocks()):
        assert nu[bi].shape == b.pressure.shape
        n = b.pressure.numel()
        assert_close(nu[bi], ref_nu[b.globalOffset:b.globalOffset + n].reshape(b.pressure.shape), tol,
                     "block %d SGS viscosity" % bi)


def test_sgs_coefficient_must_be_a_host_scalar():
    g = scenarios.cavity(dtype=torch.float32, device="cuda")
    with pytest.raises(RuntimeError):
        P.SGSviscosityIncompressibleSmagorinsky(g, torch.tensor([0.1], dtype=torch.float32, device="cuda"))
    with pytest.raises(RuntimeError):
        P.SGSviscosityIncompressibleSmagorinsky(g, torch.tensor([0.1, 0.2], dtype=torch.float32))
