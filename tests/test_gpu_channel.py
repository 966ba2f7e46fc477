"""3-D channel with a CNN sub-grid forcing (the 3-D channel configuration): before every step the block velocity
source is recomputed from the velocities (a static streamwise forcing plus a random-init three-layer Conv3d model),
and a checkpointed rollout (checkpoint.rollout_with_sources) is differentiated end to end.
- Reduced size, FP64: the gradients of the CNN weights match central finite differences of the loss.
- Full size (192 x 96 x 96 cells), FP32: two checkpointed steps forward and backward, every solve converged, finite
  gradients for every CNN parameter, velocities that stay finite and close to divergence-free."""
import pytest
import torch

from pict_amd import PISOtorch as P
from pict_amd import PISOtorch_simulation as S
from pict_amd import checkpoint as CK
from pict_amd import configs

pytestmark = pytest.mark.gpu


def _setup(dtype, size, channels, tol):
    dom, dt, forcing = configs.channel_c5(*size, dtype=dtype, device="cuda")
    sim = S.Simulation(dom, time_step=dt, corrector_steps=2, non_orthogonal=False, advection_tol=tol,
                       pressure_tol=tol, advect_passive_scalar=False)
    net = configs.sgs_cnn(channels, seed=0, dtype=dtype)
    return dom, sim, net, forcing


def _loss(dtype, size, channels, steps, tol, net_state=None, target=None):
    dom, sim, net, forcing = _setup(dtype, size, channels, tol)
    if net_state is not None:
        net.load_state_dict(net_state)
    g = forcing.reshape(1, 3, 1, 1, 1)
    out = CK.rollout_with_sources(sim, steps, lambda u: [g + 0.05 * net(u[0])])
    if target is None:
        target = torch.zeros_like(out[0])
    return ((out[0] - target) ** 2).mean(), net, dom, out


def test_channel_cnn_gradients_match_finite_differences():
    dtype, size, ch, steps, tol = torch.float64, (16, 12, 8), 4, 3, 1e-12
    loss, net, _, _ = _loss(dtype, size, ch, steps, tol)
    loss.backward()
    params = dict(net.named_parameters())
    state = {k: v.detach().clone() for k, v in net.state_dict().items()}
    checks = [("0.weight", (1, 2, 0, 1, 2)), ("2.weight", (3, 1, 2, 2, 0)), ("4.bias", (1,)), ("4.weight", (0, 2, 1, 0, 1))]
    eps = 1e-5
    for name, idx in checks:
        grad = float(params[name].grad[idx])
        vals = []
        for sgn in (1, -1):
            st = {k: v.clone() for k, v in state.items()}
            st[name][idx] += sgn * eps
            l, _, _, _ = _loss(dtype, size, ch, steps, tol, st)
            vals.append(float(l))
        fd = (vals[0] - vals[1]) / (2 * eps)
        rel = abs(fd - grad) / max(abs(fd), 1e-12)
        print(" %s%s: autograd %.6e, central difference %.6e, rel %.2e" % (name, idx, grad, fd, rel), flush=True)
        assert rel < 1e-4


def test_channel_full_size_forward_backward():
    dtype, size, ch, steps, tol = torch.float32, (192, 96, 96), 16, 2, 1e-6
    d0, _, _, _ = _setup(dtype, size, 1, tol)
    div0 = float(P.ComputeVelocityDivergence(d0).abs().max())
    del d0
    loss, net, dom, out = _loss(dtype, size, ch, steps, tol)
    loss.backward()
    torch.cuda.synchronize()
    print(" full-size channel: loss %.6e" % float(loss), flush=True)
    for name, p in net.named_parameters():
        assert p.grad is not None and torch.isfinite(p.grad).all(), name
    assert any(float(p.grad.abs().max()) > 0 for p in net.parameters())
    assert torch.isfinite(out[0]).all()
    div = float(P.ComputeVelocityDivergence(dom).abs().max())
    print(" max |div u| %.3e (initial state %.3e), max |u| %.3e" % (div, div0, float(out[0].abs().max())),
          flush=True)
    # the projection removes the initial noise's divergence (the reference's step check: the divergence falls,
    # tests/pisosteps.py:206-250)
    assert div < div0
