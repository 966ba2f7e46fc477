"""PISOtorch.matmul / matmulGrad / VectorToDiagMatrix / InvertMatrix (HIP, pict_matvec.hip) against the oracle:
every operand kind (matrix x matrix, matrix x vector, vector x matrix, vector x vector), every combination of the six
transpose/invert flags, 1-3 dimensions, FP32 (1e-5 relative) and FP64 (1e-12 relative); PISOtorch_diff.matmul's
gradients and its refusals; the in-place inverse; argument errors; and a field of the 3-D channel's size
(192 x 96 x 96 cells) checked on sampled cells."""
import itertools

import pytest
import torch

from oracle import oracle as O
from pict_amd import PISOtorch as P
from pict_amd import PISOtorch_diff as PD

pytestmark = pytest.mark.gpu

SHAPES = {1: [37], 2: [19, 23], 3: [5, 7, 9]}
TOL = {torch.float32: 1e-5, torch.float64: 1e-12}


def _rand(D, channels, seed, dtype, shape=None, well_conditioned=False):
    g = torch.Generator().manual_seed(seed)
    t = torch.rand([1, channels] + (shape or SHAPES[D]), generator=g, dtype=torch.float64) * 2 - 1
    if well_conditioned and channels == D * D:
        for k in range(D):
            t[0, k * D + k] += 3.0
    return t.to(dtype).cuda()


def _close(got, ref, dtype, what):
    ref = ref.to(torch.float64)
    err = float((got.double().cpu() - ref).abs().max()) / max(float(ref.abs().max()), 1e-30)
    assert err < TOL[dtype], (what, err)


def _operands(D, kind, dtype, seed):
    am, bm = kind[0] == "m", kind[1] == "m"
    A = _rand(D, D * D if am else D, seed, dtype, well_conditioned=True)
    B = _rand(D, D * D if bm else D, seed + 1, dtype, well_conditioned=True)
    out_ch = D * D if (am and bm) else D if (am or bm) else 1
    return A, B, out_ch


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("D", [1, 2, 3])
@pytest.mark.parametrize("kind", ["mm", "mv", "vm", "vv"])
def test_matmul_and_grad_every_flag(dtype, D, kind):
    if D == 1 and kind != "mm":
        pytest.skip("with one dimension every operand is a 1x1 matrix")
    print("case", dtype, D, kind, flush=True)
    A, B, out_ch = _operands(D, kind, dtype, 3)
    G = _rand(D, out_ch, 9, dtype)
    Ac, Bc, Gc = A.cpu().double(), B.cpu().double(), G.cpu().double()
    for flags in itertools.product([False, True], repeat=6):
        out = P.matmul(A, B, *flags)
        assert out.shape == (1, out_ch) + tuple(SHAPES[D]) and out.dtype == dtype
        _close(out, O.matmul(Ac, Bc, *flags), dtype, ("out", flags))
        ga, gb = P.matmulGrad(A, B, G, *flags)
        ra, rb = O.matmulGrad(Ac, Bc, Gc, *flags)
        _close(ga, ra, dtype, ("gradA", flags))
        _close(gb, rb, dtype, ("gradB", flags))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("D", [1, 2, 3])
def test_diag_and_inverse(dtype, D):
    v = _rand(D, D, 4, dtype)
    _close(P.VectorToDiagMatrix(v), O.VectorToDiagMatrix(v.cpu().double()), dtype, "diag")
    m = _rand(D, D * D, 5, dtype, well_conditioned=True)
    inv = P.InvertMatrix(m)
    _close(inv, O.InvertMatrix(m.cpu().double()), dtype, "inverse")
    # in place: the same tensor, now holding the inverse
    m2 = m.clone()
    r = P.InvertMatrix(m2, inPlace=True)
    assert r.data_ptr() == m2.data_ptr()
    assert torch.equal(r, inv)


def test_diff_matmul_gradients_and_refusals():
    D, dtype = 3, torch.float64
    A = _rand(D, 9, 6, dtype, well_conditioned=True).requires_grad_(True)
    v = _rand(D, 3, 7, dtype).requires_grad_(True)
    G = _rand(D, 3, 8, dtype)
    out = PD.matmul(A, v, transposeA=True)
    (out * G).sum().backward()
    ra, rv = O.matmulGrad(A.detach().cpu(), v.detach().cpu(), G.cpu(), transposeA=True)
    _close(A.grad, ra, dtype, "A.grad")
    _close(v.grad, rv, dtype, "v.grad")
    # gradients through an inverted operand are refused, as in the reference
    B = _rand(D, 9, 9, dtype, well_conditioned=True).requires_grad_(True)
    with pytest.raises(NotImplementedError):
        PD.matmul(B, v, invertA=True).sum().backward()
    # ... but not when that operand needs none
    w = _rand(D, 3, 10, dtype).requires_grad_(True)
    PD.matmul(B.detach(), w, invertA=True).sum().backward()
    assert w.grad is not None


def test_argument_errors():
    v2 = _rand(2, 2, 1, torch.float32)
    with pytest.raises(RuntimeError):
        P.matmul(v2, _rand(3, 3, 2, torch.float32))  # dimensionality differs
    with pytest.raises(RuntimeError):
        P.matmul(_rand(2, 3, 1, torch.float32), v2)  # 3 channels in 2-D: neither vector nor matrix
    with pytest.raises(RuntimeError):
        P.matmul(v2, _rand(2, 2, 2, torch.float64))  # dtype differs
    with pytest.raises(RuntimeError):
        P.matmulGrad(v2, v2, _rand(2, 2, 3, torch.float32))  # output gradient of a vector product is 1 channel
    with pytest.raises(RuntimeError):
        P.InvertMatrix(v2)
    with pytest.raises(RuntimeError):
        P.VectorToDiagMatrix(_rand(2, 4, 1, torch.float32))


def test_channel_sized_field():
    """Matrix x vector on the 3-D channel's grid (192 x 96 x 96 cells), FP32, plain, transposed and transposed +
    inverted matrix; sampled cells against the oracle."""
    D, shape, dtype = 3, [96, 96, 192], torch.float32
    M = _rand(D, 9, 11, dtype, shape, True)
    v = _rand(D, 3, 12, dtype, shape)
    g = torch.Generator().manual_seed(13)
    idx = torch.randint(0, 96 * 96 * 192, (4096,), generator=g)
    # the sampled cells as a 1 x 1 x 4096 grid of the same dimensionality
    Ms = M.reshape(9, -1)[:, idx].cpu().double().reshape(1, 9, 1, 1, 4096)
    vs = v.reshape(3, -1)[:, idx].cpu().double().reshape(1, 3, 1, 1, 4096)
    for flags in [(False,) * 6, (True, False, False, False, False, False), (True, True, False, False, False, False)]:
        out = P.matmul(M, v, *flags)
        torch.cuda.synchronize()
        _close(out.reshape(3, -1)[:, idx].reshape(1, 3, 1, 1, 4096), O.matmul(Ms, vs, *flags), dtype, flags)
