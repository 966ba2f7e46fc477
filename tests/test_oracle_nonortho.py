"""Pins of the oracle's non-orthogonal terms (matrix entries, explicit RHS terms, corner walks) by checks that do
not restate them.

The reference splits the Laplacian of a non-orthogonal grid between the matrix and an explicit right-hand side
(PISO_multiblock_cuda_kernel.cu:3738-3800, :4877-4932, :3033-3190, :5455-5477): with the driver's flags 25 the
cross-derivative terms through the centre cell and its face neighbours go into the matrix, the ones through the
diagonal cells are evaluated from the previous iterate.  The operator the two parts form together,

    pressure:  L p = P p - N(p)       (N: what SetupPressureRHSdiv adds to the divergence)
    velocity:  det * (C u - u / dt) - det * rhs(u)   (rhs: SetupAdvectionVelocity's output with a zero old
               velocity, whose only term is then the explicit non-orthogonal one)

is fixed by the mathematics:
- on an affine (sheared) grid the finite-volume Laplacian is exact for quadratic fields away from boundaries:
  L p = det * (p_xx + p_yy [+ p_zz]) (pressure, A = 1) and -nu * det * Laplacian(u) (velocity);
- the operator does not depend on how the flags split it: CENTER|DIRECT matrix + DIAGONAL rhs (25) and
  CENTER matrix + DIRECT|DIAGONAL rhs (17|2|8) give the same L for any field (cells whose cross coefficients the
  reference interpolates the same way in both, i.e. not the end cells of an axis with a wall there);
- on an orthogonal grid every non-orthogonal term vanishes;
- 1/A enters the pressure operator linearly.
Mutations of the oracle (a dropped 1/2 in the face interpolation, a flipped tangential sign, a wrong corner count)
fail these checks (tests/test_oracle_mutations.py).
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from scenarios import sheared_box, sheared_coords

FLAGS_MATRIX_SPLIT = 1 | 16 | 8   # DIRECT_MATRIX | CENTER_MATRIX | DIAGONAL_RHS (the driver's flags)
FLAGS_RHS_SPLIT = 16 | 2 | 8      # CENTER_MATRIX | DIRECT_RHS | DIAGONAL_RHS


def cell_centres(*coords):
    """Physical cell centres [N, D] (mean of each cell's vertices) of blocks with these vertex coordinates, in
    global cell order."""
    out = []
    for c in coords:
        D = c.size(1)
        if D == 2:
            cc = 0.25 * (c[..., :-1, :-1] + c[..., 1:, :-1] + c[..., :-1, 1:] + c[..., 1:, 1:])
        else:
            cc = 0.125 * sum(c[..., dz:c.size(2) - 1 + dz, dy:c.size(3) - 1 + dy, dx:c.size(4) - 1 + dx]
                             for dz in (0, 1) for dy in (0, 1) for dx in (0, 1))
        out.append(cc[0].reshape(D, -1).T)
    return torch.cat(out).double()


def dets(domain):
    return torch.cat([b.transform[..., -1].reshape(-1) for b in domain.blocks]).double()


def quadratic(x, seed=0):
    """Random quadratic q(x) = x^T H x / 2 + g.x + c; returns (values, Laplacian)."""
    D = x.size(1)
    g = torch.Generator().manual_seed(seed)
    H = torch.rand(D, D, generator=g, dtype=torch.float64) * 2 - 1
    H = H + H.T
    gv = torch.rand(D, generator=g, dtype=torch.float64)
    q = 0.5 * ((x @ H) * x).sum(1) + x @ gv + 0.3
    return q, float(torch.trace(H))


def csr_apply(A, v):
    sp = O.csr_to_scipy(A)
    return torch.from_numpy(sp @ v.numpy())


def interior_mask(domain, margin=1):
    m = []
    for b in domain.blocks:
        sz = [b.getSizes().x, b.getSizes().y, b.getSizes().z][:domain.getSpatialDims()]
        grids = torch.meshgrid(*[torch.arange(n) for n in reversed(sz)], indexing="ij")
        ok = torch.ones_like(grids[0], dtype=torch.bool)
        for ax, n in enumerate(reversed(sz)):
            ok &= (grids[ax] >= margin) & (grids[ax] <= n - 1 - margin)
        m.append(ok.reshape(-1))
    return torch.cat(m)


def pressure_operator(dom, p, flags):
    """L p = P p - N(p) with the pressure matrix and the explicit RHS terms of `flags`."""
    dt = torch.tensor([1.0], dtype=torch.float64)
    O.SetupPressureMatrix(dom, dt, flags)
    dom.setPressureResult(p.clone())
    dom.setPressureRHSdiv(torch.zeros_like(p))
    O.PressureRHSaddNonOrtho(dom, flags)
    return csr_apply(dom.P, p) - dom.pressureRHSdiv


def velocity_operator(dom, u, flags, dt=0.1):
    """det (C u - u/dt) - det rhs(u) with a zero old block velocity (C linearised on u = 0: diffusion only)."""
    for b in dom.blocks:
        b.setVelocity(torch.zeros_like(b.velocity))
    dom.UpdateDomainData()
    t = torch.tensor([dt], dtype=torch.float64)
    O.SetupAdvectionMatrix(dom, t, flags)
    dom.setVelocityResult(u.clone())
    O.SetupAdvectionVelocity(dom, t, flags, False)
    N = dom.getTotalSize()
    D = dom.getSpatialDims()
    det = dets(dom)
    out = []
    for k in range(D):
        uk = u[k * N:(k + 1) * N]
        out.append(det * (csr_apply(dom.C, uk) - uk / dt) - det * dom.velocityRHS[k * N:(k + 1) * N])
    return torch.stack(out)


CASES = [((10, 9), (0.45,)), ((9, 10), (-0.3,)), ((6, 5, 7), (0.3, -0.25, 0.2))]


@pytest.mark.parametrize("size,shear", CASES)
def test_pressure_operator_exact_for_quadratics(size, shear):
    dom = sheared_box(device="cpu", size=size, shear=shear, walls_moving=False)
    dom.setA(torch.ones_like(dom.A))
    x = cell_centres(sheared_coords(size, shear))
    q, lap = quadratic(x, seed=len(size))
    L = pressure_operator(dom, q, FLAGS_MATRIX_SPLIT)
    m = interior_mask(dom)
    want = dets(dom) * lap
    err = (L - want)[m].abs().max().item() / abs(want).max().item()
    assert err < 1e-10, err
    # without the non-orthogonal terms the sheared grid is visibly wrong: the check has teeth
    L0 = pressure_operator(dom, q, 0)
    assert (L0 - want)[m].abs().max().item() / abs(want).max().item() > 1e-2


@pytest.mark.parametrize("size,shear", CASES)
def test_pressure_operator_scales_with_inverse_A(size, shear):
    dom = sheared_box(device="cpu", size=size, shear=shear, walls_moving=False)
    x = cell_centres(sheared_coords(size, shear))
    q, lap = quadratic(x, seed=5)
    dom.setA(torch.full_like(dom.A, 4.0))
    L = pressure_operator(dom, q, FLAGS_MATRIX_SPLIT)
    m = interior_mask(dom)
    want = dets(dom) * lap / 4.0
    assert (L - want)[m].abs().max().item() / abs(want).max().item() < 1e-10


@pytest.mark.parametrize("size,shear", CASES)
def test_velocity_operator_exact_for_quadratics(size, shear):
    dom = sheared_box(device="cpu", size=size, shear=shear, walls_moving=False)
    nu = float(dom.viscosity[0])
    x = cell_centres(sheared_coords(size, shear))
    D = len(size)
    qs, laps = zip(*[quadratic(x, seed=10 + k) for k in range(D)])
    u = torch.cat(qs)
    L = velocity_operator(dom, u, FLAGS_MATRIX_SPLIT)
    m = interior_mask(dom)
    for k in range(D):
        want = -nu * dets(dom) * laps[k]
        err = (L[k] - want)[m].abs().max().item() / abs(want).max().item()
        assert err < 1e-9, (k, err)


@pytest.mark.parametrize("periodic", [("x",), ("x", "y")])
def test_flag_split_gives_one_operator(periodic):
    """Random fields on a bent sheared grid: the same operator whichever way the flags split it."""
    dom = sheared_box(device="cpu", size=(9, 8), shear=(0.4,), periodic=periodic, bend=0.3, walls_moving=False)
    g = torch.Generator().manual_seed(3)
    dom.setA(torch.rand(dom.A.shape, generator=g, dtype=torch.float64) + 0.5)
    p = torch.rand(dom.getTotalSize(), generator=g, dtype=torch.float64)
    La = pressure_operator(dom, p, FLAGS_MATRIX_SPLIT)
    Lb = pressure_operator(dom, p, FLAGS_RHS_SPLIT)
    m = interior_mask(dom) if len(periodic) < 2 else torch.ones_like(La, dtype=torch.bool)
    assert (La - Lb)[m].abs().max().item() < 1e-12 * La.abs().max().item()
    u = torch.rand(2 * dom.getTotalSize(), generator=g, dtype=torch.float64)
    Va = velocity_operator(dom, u, FLAGS_MATRIX_SPLIT)
    Vb = velocity_operator(dom, u, FLAGS_RHS_SPLIT)
    assert (Va - Vb)[:, m].abs().max().item() < 1e-12 * Va.abs().max().item()


def test_orthogonal_grid_has_no_cross_terms():
    """On a rectangular grid every cross coefficient is zero: flags change nothing."""
    dom = sheared_box(device="cpu", size=(8, 7), shear=(0.0,), bend=0.0)
    g = torch.Generator().manual_seed(4)
    p = torch.rand(dom.getTotalSize(), generator=g, dtype=torch.float64)
    dom.setA(torch.rand(dom.A.shape, generator=g, dtype=torch.float64) + 0.5)
    assert torch.equal(pressure_operator(dom, p, FLAGS_MATRIX_SPLIT), pressure_operator(dom, p, 0))
    u = torch.rand(2 * dom.getTotalSize(), generator=g, dtype=torch.float64)
    assert torch.equal(velocity_operator(dom, u, FLAGS_MATRIX_SPLIT), velocity_operator(dom, u, 0))


def test_sheared_coords_are_affine():
    c = sheared_coords((4, 3), (0.5,), h=0.2)
    assert torch.allclose(c[0, 0, 1, 0] - c[0, 0, 0, 0], torch.tensor(0.1, dtype=torch.float64))
    assert torch.allclose(c[0, 1, 1, 0] - c[0, 1, 0, 0], torch.tensor(0.2, dtype=torch.float64))
