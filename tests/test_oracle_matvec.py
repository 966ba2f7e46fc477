"""Pins the oracle's per-cell vector/matrix operations (oracle.matmul, matmulGrad, VectorToDiagMatrix, InvertMatrix)
on the CPU, with checks that do not restate its formulas:
- the inverse against numpy.linalg.inv;
- every flag combination of matrix x matrix against torch.linalg.inv / torch.matmul compositions, and the
  identities (A B)^T = B^T A^T, (A B)^-1 = B^-1 A^-1, A^T v = v^T A, v^T A = A^T v;
- the gradients against torch.autograd of those compositions and against central finite differences, and the
  reference's zero gradients for inverted operands (matrix_vector_ops_grads.cu:19, :29, :50)."""
import itertools

import numpy as np
import pytest
import torch

from oracle import oracle as O

SHAPES = {1: [7], 2: [5, 4], 3: [3, 4, 2]}


def _rand(D, channels, seed, well_conditioned=False):
    g = torch.Generator().manual_seed(seed)
    t = torch.rand([1, channels] + SHAPES[D], generator=g, dtype=torch.float64) * 2 - 1
    if well_conditioned and channels == D * D:
        for k in range(D):
            t[0, k * D + k] += 3.0  # diagonally dominant: invertible, moderate condition number
    return t


def _as_mats(t, D):
    """[1, D*D, ...] -> [cells, D, D] (row-major channels)."""
    return t.reshape(D * D, -1).T.reshape(-1, D, D)


def _as_vecs(t, D):
    return t.reshape(D, -1).T


def _torch_product(A, B, D, am, bm, flags):
    tA, iA, tB, iB, tO, iO = flags
    a = _as_mats(A, D) if am else _as_vecs(A, D)
    b = _as_mats(B, D) if bm else _as_vecs(B, D)
    if am:
        a = a.transpose(1, 2) if tA else a
        a = torch.linalg.inv(a) if iA else a
    if bm:
        b = b.transpose(1, 2) if tB else b
        b = torch.linalg.inv(b) if iB else b
    if am and bm:
        c = a @ b
        c = torch.linalg.inv(c) if iO else c
        c = c.transpose(1, 2) if tO else c
        return c.reshape(-1, D * D).T.reshape([1, D * D] + list(A.shape[2:]))
    if am:
        c = (a @ b.unsqueeze(-1)).squeeze(-1)
    elif bm:
        c = (a.unsqueeze(1) @ b).squeeze(1)
    else:
        c = (a * b).sum(-1, keepdim=True)
    return c.T.reshape([1, c.shape[1]] + list(A.shape[2:]))


@pytest.mark.parametrize("D", [1, 2, 3])
def test_inverse_matches_numpy(D):
    m = _rand(D, D * D, 1, True)
    inv = O.InvertMatrix(m)
    ref = np.linalg.inv(_as_mats(m, D).numpy())
    assert np.abs(_as_mats(inv, D).numpy() - ref).max() < 1e-13
    # identity: A A^-1 = I
    prod = _as_mats(O.matmul(m, inv), D).numpy()
    assert np.abs(prod - np.eye(D)).max() < 1e-13


@pytest.mark.parametrize("D", [1, 2, 3])
def test_matrix_products_every_flag(D):
    A, B = _rand(D, D * D, 2, True), _rand(D, D * D, 3, True)
    for flags in itertools.product([False, True], repeat=6):
        got = O.matmul(A, B, *flags)
        ref = _torch_product(A, B, D, True, True, flags)
        assert torch.allclose(got, ref, rtol=1e-11, atol=1e-12), flags
    # identities that a swapped operand or flag would break
    assert torch.allclose(O.matmul(A, B, transposeOutput=True), O.matmul(B, A, transposeA=True, transposeB=True))
    assert torch.allclose(O.matmul(A, B, invertOutput=True), O.matmul(B, A, invertA=True, invertB=True))


@pytest.mark.parametrize("D", [1, 2, 3])
def test_vector_products(D):
    M = _rand(D, D * D, 4, True)
    v, w = _rand(D, D, 5), _rand(D, D, 6)
    if D > 1:
        for flags in itertools.product([False, True], repeat=6):
            assert torch.allclose(O.matmul(M, v, *flags), _torch_product(M, v, D, True, False, flags))
            assert torch.allclose(O.matmul(v, M, *flags), _torch_product(v, M, D, False, True, flags))
        # A^T v = v^T A; the output flags do not act on vectors
        assert torch.allclose(O.matmul(M, v, transposeA=True), O.matmul(v, M))
        assert torch.allclose(O.matmul(M, v, transposeOutput=True, invertOutput=True), O.matmul(M, v))
        assert torch.allclose(O.matmul(v, w, True, True, True, True, True, True), (v * w).sum(1, keepdim=True))
    diag = O.VectorToDiagMatrix(v)
    assert torch.allclose(_as_mats(diag, D), torch.diag_embed(_as_vecs(v, D)))


def _autograd(A, B, G, D, am, bm, flags):
    a = A.clone().requires_grad_(True)
    b = B.clone().requires_grad_(True)
    out = _torch_product(a, b, D, am, bm, flags)
    ga, gb = torch.autograd.grad((out * G).sum(), [a, b], allow_unused=True)
    return (torch.zeros_like(A) if ga is None else ga), (torch.zeros_like(B) if gb is None else gb)


@pytest.mark.parametrize("D", [1, 2, 3])
@pytest.mark.parametrize("kind", ["mm", "mv", "vm", "vv"])
def test_gradients_match_autograd(D, kind):
    am, bm = kind[0] == "m", kind[1] == "m"
    A = _rand(D, D * D if am else D, 7, True)
    B = _rand(D, D * D if bm else D, 8, True)
    out_ch = D * D if (am and bm) else D if (am or bm) else 1
    if D == 1 and kind != "mm":
        pytest.skip("with one dimension every operand is a 1x1 matrix")
    G = _rand(D, out_ch, 9)
    for flags in itertools.product([False, True], repeat=6):
        tA, iA, tB, iB, tO, iO = flags
        ga, gb = O.matmulGrad(A, B, G, *flags)
        ra, rb = _autograd(A, B, G, D, am, bm, flags)
        # the reference gives no gradient to an inverted operand, and none at all for an inverted matrix product
        # or one with both operands inverted
        if am and bm and (iO or (iA and iB)):
            ra, rb = torch.zeros_like(A), torch.zeros_like(B)
        if am and iA:
            ra = torch.zeros_like(A)
        if bm and iB:
            rb = torch.zeros_like(B)
        assert torch.allclose(ga, ra, rtol=1e-10, atol=1e-11), (kind, flags)
        assert torch.allclose(gb, rb, rtol=1e-10, atol=1e-11), (kind, flags)


def test_gradient_finite_differences():
    """One flag set of matrix x matrix, checked against central differences of the oracle's own product."""
    D = 3
    A, B, G = _rand(D, 9, 10, True), _rand(D, 9, 11, True), _rand(D, 9, 12)
    flags = (True, False, False, True, True, False)
    ga, _ = O.matmulGrad(A, B, G, *flags)
    eps = 1e-6
    for idx in [(0, 0, 1, 2, 0), (0, 5, 2, 3, 1), (0, 8, 0, 0, 1)]:
        Ap, Am = A.clone(), A.clone()
        Ap[idx] += eps
        Am[idx] -= eps
        fd = ((O.matmul(Ap, B, *flags) - O.matmul(Am, B, *flags)) * G).sum() / (2 * eps)
        assert abs(float(fd) - float(ga[idx])) < 1e-7
