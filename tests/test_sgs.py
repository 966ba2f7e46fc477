"""sgs.Conv3dSame: the same outputs and gradients as torch.nn.Conv3d(in, out, 3, padding=1) with its own backward
(input gradient by the flipped-kernel convolution, weight gradient by the chunked batched product), on the CPU in
FP64, for channel counts of the SGS model and a field larger than one chunk."""
import pytest
import torch

from pict_amd import sgs


@pytest.mark.parametrize("cin,cout,shape", [(3, 4, (5, 6, 7)), (4, 4, (9, 8, 130)), (4, 3, (3, 3, 3))])
def test_conv3d_same_matches_reference(cin, cout, shape):
    torch.manual_seed(0)
    ref = torch.nn.Conv3d(cin, cout, 3, padding=1).double()
    mine = sgs.Conv3dSame(cin, cout).double()
    mine.load_state_dict(ref.state_dict())
    x = torch.randn((1, cin) + shape, dtype=torch.float64)
    g = torch.randn((1, cout) + shape, dtype=torch.float64)
    xr, xm = x.clone().requires_grad_(True), x.clone().requires_grad_(True)
    yr, ym = ref(xr), mine(xm)
    assert torch.allclose(yr, ym, rtol=1e-12, atol=1e-12)
    (yr * g).sum().backward()
    (ym * g).sum().backward()
    assert torch.allclose(xr.grad, xm.grad, rtol=1e-11, atol=1e-11)
    assert torch.allclose(ref.weight.grad, mine.weight.grad, rtol=1e-11, atol=1e-11)
    assert torch.allclose(ref.bias.grad, mine.bias.grad, rtol=1e-11, atol=1e-11)
