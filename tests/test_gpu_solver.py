"""Krylov kernels on systems built here: shifted 5-point (2-D) and 7-point (3-D) periodic Laplacians, the
singular periodic Laplacian, non-symmetric values on a symmetric pattern, and irregular patterns with long rows.

The register-resident CG (one right-hand side) is checked for every rows-per-thread instance, both reduction
layouts (norm-2 criteria and absolute criteria), the best-result and zero right-hand-side paths, transposed
solves and rows read from memory:
- FP64 solutions against a direct FP64 solve (scipy.sparse.linalg.spsolve) at 1e-8 relative;
- FP32 and FP64 against the grid-wide single-reduction CG kernel (the same iterates in exact arithmetic):
  iteration counts within 3% + 2, solutions within the solver tolerance's reach.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

from pict_amd import PISOtorch as P
from pict_amd import _abi

pytestmark = pytest.mark.gpu

NORM2N = P.ConvergenceCriterion.NORM2_NORMALIZED
ABS_MAX = P.ConvergenceCriterion.ABS_MAX


def periodic_laplacian(shape, shift):
    """-(periodic 2*D+1 point Laplacian) + shift*I on a grid of `shape` (x fastest), as scipy CSR."""
    n = int(np.prod(shape))
    idx = np.arange(n).reshape(shape[::-1])
    rows, cols, vals = [idx.ravel()], [idx.ravel()], [np.full(n, 2.0 * len(shape) + shift)]
    for ax in range(len(shape)):
        for s in (-1, 1):
            rows.append(idx.ravel())
            cols.append(np.roll(idx, s, axis=ax).ravel())
            vals.append(np.full(n, -1.0))
    m = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    m.sum_duplicates()
    m.sort_indices()
    return m


def irregular(n, seed, long_rows=8, long_len=12):
    """SPD matrix on a random symmetric pattern: a banded core plus `long_rows` rows/columns of `long_len`
    far-away entries (rows longer than the register cache, columns owned by other workgroups)."""
    rng = np.random.default_rng(seed)
    m = periodic_laplacian((n,), 0.5).tolil()
    for r in rng.choice(n, long_rows, replace=False):
        for c in rng.choice(n, long_len, replace=False):
            if c != r:
                m[r, c] = m[c, r] = -0.05
    m = m.tocsr()
    m = m + sp.diags(np.asarray(abs(m).sum(axis=1)).ravel() * 0.1)  # diagonally dominant
    m.sort_indices()
    return m


def to_pict(m, dtype):
    return P.CSRmatrix(torch.tensor(m.data, dtype=dtype).cuda(), torch.tensor(m.indices, dtype=torch.int32).cuda(),
                       torch.tensor(m.indptr, dtype=torch.int32).cuda())


def solve(A, b, tol=1e-10, crit=NORM2N, maxit=5000, transpose=False, best=False, x0=None):
    x = torch.zeros_like(b) if x0 is None else x0.clone()
    info = P.SolveLinear(A, b.contiguous(), x, maxit, tol, crit, False, False, 0, transpose, False, best, False)
    return x, info[0]


def kernel_name(A, crit=NORM2N):
    o = P._solve_opts(5000, 1e-10, crit, False, False, 0, False, False, False, False)
    return P.solver_kernel_name(A, 1, o)


def rhs(n, seed, dtype, zero_mean=False):
    g = torch.Generator().manual_seed(seed)
    b = torch.rand(n, generator=g, dtype=torch.float64) * 2 - 1
    if zero_mean:
        b -= b.mean()
    return b.to(dtype).cuda()


# (grid, workgroup budget) -> rows per thread of the kernel chosen
SIZES = [((64, 64), 64, 1), ((128, 128), 32, 2), ((256, 256), 64, 4), ((256, 256), 32, 8), ((256, 128), 64, 2),
         ((32, 16, 24), 64, 1), ((48, 32, 32), 48, 4)]


@pytest.mark.parametrize("grid,wgs,rpt", SIZES)
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_local_cg_matches_direct_and_grid_kernel(grid, wgs, rpt, dtype, monkeypatch):
    if dtype == torch.float64 and rpt > 4:
        pytest.skip("FP64 rows use at most 4 rows per thread")
    print("case", grid, wgs, rpt, dtype, flush=True)
    m = periodic_laplacian(grid, 0.05)
    A = to_pict(m, dtype)
    b = rhs(m.shape[0], 3, dtype)
    monkeypatch.setenv("PICT_CG_LOCAL_WGS", str(wgs))
    name = kernel_name(A)
    nz = 5 if len(grid) == 2 else 7
    t = "float" if dtype == torch.float32 else "double"
    assert name == "pict::k_cg_local<%s, %d, %d, false>" % (t, rpt, nz), name
    tol = 1e-12 if dtype == torch.float64 else 1e-6
    x, info = solve(A, b, tol)
    assert info.converged and info.isFiniteResidual
    monkeypatch.setenv("PICT_CG_LOCAL_WGS", "0")
    assert "k_cg1" in kernel_name(A)
    xr, ir = solve(A, b, tol)
    print(" its local %d grid-wide %d" % (info.usedIterations, ir.usedIterations), flush=True)
    assert abs(info.usedIterations - ir.usedIterations) <= 2 + 0.03 * ir.usedIterations
    if dtype == torch.float64:
        xd = spla.spsolve(m.tocsc(), b.cpu().numpy())
        err = np.abs(x.cpu().numpy() - xd).max() / np.abs(xd).max()
        assert err < 1e-8, err
        assert (x - xr).abs().max() / xr.abs().max() < 1e-8
    else:
        assert float((x - xr).abs().max() / xr.abs().max()) < 1e-3


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_local_cg_absolute_criterion_and_best_result(dtype):
    """ABS_MAX needs the sum/maximum slots; returnBestResult keeps the iterate of the smallest criterion."""
    m = periodic_laplacian((128, 64), 0.02)
    A = to_pict(m, dtype)
    b = rhs(m.shape[0], 7, dtype)
    assert kernel_name(A, ABS_MAX).endswith(", true>")
    tol = 1e-9 if dtype == torch.float64 else 1e-4
    x, info = solve(A, b, tol, crit=ABS_MAX)
    assert info.converged
    r = b.double().cpu().numpy() - m @ x.double().cpu().numpy()
    assert np.abs(r).max() < 10 * tol
    # too few iterations: not converged, the best iterate is returned and its criterion reported
    xb, ib = solve(A, b, 1e-30, maxit=40, best=True)
    assert not ib.converged and ib.usedIterations <= 40
    rb = b.double().cpu().numpy() - m @ xb.double().cpu().numpy()
    crit = float(np.linalg.norm(rb) / np.sqrt(m.shape[0]))
    assert abs(crit - ib.finalResidual) <= 1e-3 * ib.finalResidual + (1e-6 if dtype == torch.float32 else 1e-12)


def test_local_cg_zero_rhs():
    m = periodic_laplacian((64, 64), 0.1)
    A = to_pict(m, torch.float32)
    b = torch.zeros(m.shape[0], dtype=torch.float32, device="cuda")
    x0 = rhs(m.shape[0], 1, torch.float32)
    # the shortcut (PISOtorch_diff's adjoint solves): x = 0, no iterations
    x = x0.clone()
    o = P._solve_opts(100, 1e-6, NORM2N, False, False, 0, False, False, False, False)
    o.zero_rhs_shortcut = 1
    infos = P._solve(A, b, x, o)
    assert infos[0].converged and infos[0].usedIterations == -1 and float(x.abs().max()) == 0.0
    # without it, CG from x0 drives x to 0
    x, info = solve(A, b, 1e-7, x0=x0)
    assert info.converged and float(x.abs().max()) < 1e-5


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_local_cg_transposed_and_singular(dtype):
    """Transposed solve of a non-symmetric matrix with a symmetric pattern (the adjoint solves), and the singular
    periodic Laplacian with a zero-mean right-hand side (the pressure solve)."""
    m = periodic_laplacian((96, 64), 0.3)
    rng = np.random.default_rng(5)
    m.data = m.data * (1 + 0.1 * rng.random(m.data.shape))  # values no longer symmetric
    A = to_pict(m, dtype)
    b = rhs(m.shape[0], 2, dtype)
    # CG on a non-symmetric matrix still runs its recurrences; the transposed solve must give the solve of m.T
    # that the same CG gives on an explicit transpose
    mt = m.T.tocsr()
    mt.sort_indices()
    At = to_pict(mt, dtype)
    tol = 1e-10 if dtype == torch.float64 else 1e-6
    xt, it = solve(A, b, tol, transpose=True, maxit=300)
    xe, ie = solve(At, b, tol, maxit=300)
    assert it.usedIterations == ie.usedIterations
    assert float((xt - xe).abs().max()) == 0.0  # same values, same order: bitwise
    ms = periodic_laplacian((64, 96), 0.0)
    As = to_pict(ms, dtype)
    bs = rhs(ms.shape[0], 4, dtype, zero_mean=True)
    xs, info = solve(As, bs, 1e-10 if dtype == torch.float64 else 1e-5)
    assert info.converged
    r = bs.double().cpu().numpy() - ms @ xs.double().cpu().numpy()
    assert np.abs(r).max() < (1e-8 if dtype == torch.float64 else 1e-3)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_local_cg_rows_from_memory(dtype, monkeypatch):
    """Rows longer than the register cache and rows with many foreign columns take the memory path."""
    m = irregular(20000, 3)
    A = to_pict(m, dtype)
    b = rhs(m.shape[0], 9, dtype)
    monkeypatch.setenv("PICT_CG_LOCAL_WGS", "64")
    assert "k_cg_local" in kernel_name(A)
    tol = 1e-12 if dtype == torch.float64 else 1e-6
    x, info = solve(A, b, tol)
    assert info.converged
    monkeypatch.setenv("PICT_CG_LOCAL_WGS", "0")
    xr, ir = solve(A, b, tol)
    assert abs(info.usedIterations - ir.usedIterations) <= 2 + 0.03 * ir.usedIterations
    if dtype == torch.float64:
        xd = spla.spsolve(m.tocsc(), b.cpu().numpy())
        assert np.abs(x.cpu().numpy() - xd).max() / np.abs(xd).max() < 1e-8
    else:
        assert float((x - xr).abs().max() / xr.abs().max()) < 1e-3


def test_local_cg_deterministic():
    """Every workgroup reduces the same values in the same order: repeated solves are bitwise identical."""
    m = periodic_laplacian((256, 256), 0.01)
    A = to_pict(m, torch.float32)
    b = rhs(m.shape[0], 6, torch.float32)
    x1, i1 = solve(A, b, 1e-6)
    for _ in range(3):
        x2, i2 = solve(A, b, 1e-6)
        assert i2.usedIterations == i1.usedIterations and torch.equal(x1, x2)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_streaming_cg_lean_kernel(dtype, monkeypatch):
    """A 3-D system with more rows than the grid-wide CG's register row cache holds (64 x 64 x 80 = 327,680 rows)
    runs the lean streaming kernel; the textbook two-barrier CG (the reference's iterates) agrees on the iteration
    count within 3% + 2, and the true FP64 residual |b - A x|_2 / sqrt(n) of both is at the solver tolerance."""
    print("case lean streaming CG", dtype, flush=True)
    m = periodic_laplacian((64, 64, 80), 0.05)
    n = m.shape[0]
    A = to_pict(m, dtype)
    b = rhs(n, 5, dtype)
    monkeypatch.setenv("PICT_CG_LOCAL_WGS", "0")
    t = "float" if dtype == torch.float32 else "double"
    assert kernel_name(A) == "pict::k_cg1<%s, 1, true>" % t
    tol = 1e-10 if dtype == torch.float64 else 1e-6
    x, info = solve(A, b, tol)
    assert info.converged and info.isFiniteResidual
    monkeypatch.setenv("PICT_CG_TEXTBOOK", "1")
    assert kernel_name(A) == "pict::k_cg<%s, 1>" % t
    xt, it = solve(A, b, tol)
    assert it.converged
    print(" its lean %d textbook %d" % (info.usedIterations, it.usedIterations), flush=True)
    assert abs(info.usedIterations - it.usedIterations) <= 2 + 0.03 * it.usedIterations
    bd = b.double().cpu().numpy()
    for sol in (x, xt):
        res = np.linalg.norm(bd - m @ sol.double().cpu().numpy()) / np.sqrt(n)
        print(" true residual %.3e" % res, flush=True)
        assert res < (2 * tol if dtype == torch.float64 else 1e-5)
    # a system the row cache holds keeps the cached kernel
    small = to_pict(periodic_laplacian((64, 64, 16), 0.05), dtype)
    monkeypatch.delenv("PICT_CG_TEXTBOOK")
    assert kernel_name(small) == "pict::k_cg1<%s, 1, false>" % t


@pytest.mark.parametrize("bicg", [False, True])
def test_large_grid_wide_solves_stay_resident(bicg, monkeypatch):
    """Grid-wide kernels with few registers on the 3-D channel's size (192 x 96 x 96 = 1,769,472 rows): the
    textbook CG and the one-right-hand-side BiCGStab.  Their persistent grids are sized at no more than four
    workgroups per compute unit; with the occupancy query's seven or eight, the textbook CG's barrier ran into
    its spin limit on this system.  Thirty iterations run without a barrier error and reduce the residual."""
    print("case large grid-wide", "bicgstab" if bicg else "textbook cg", flush=True)
    m = periodic_laplacian((192, 96, 96), 0.05)
    A = to_pict(m, torch.float32)
    b = rhs(m.shape[0], 7, torch.float32)
    monkeypatch.setenv("PICT_CG_LOCAL_WGS", "0")
    if not bicg:
        monkeypatch.setenv("PICT_CG_TEXTBOOK", "1")
    x = torch.zeros_like(b)
    info = P.SolveLinear(A, b, x, 30, 1e-30, NORM2N, bicg, False, 0, False, False, False, False)[0]
    assert info.isFiniteResidual and info.usedIterations >= 29
    bd = b.double().cpu().numpy()
    res = np.linalg.norm(bd - m @ x.double().cpu().numpy()) / np.linalg.norm(bd)
    print(" relative residual after 30 iterations %.3e" % res, flush=True)
    assert res < 0.5


@pytest.mark.parametrize("bicg", [False, True])
def test_small_grids_cooperative_and_plain_launch(bicg, monkeypatch):
    """Solver grids of at most one workgroup per compute unit are ordinary launches by default;
    PICT_SOLVER_PLAIN_LAUNCH=0 launches them cooperatively.  Both give bitwise the same solution and iteration
    count (the same kernel on the same grid)."""
    print("case launch kinds", "bicgstab" if bicg else "cg", flush=True)
    m = periodic_laplacian((128, 96), 0.05)
    A = to_pict(m, torch.float32)
    b = rhs(m.shape[0], 11, torch.float32)
    out = []
    for plain in ("1", "0"):
        monkeypatch.setenv("PICT_SOLVER_PLAIN_LAUNCH", plain)
        x = torch.zeros_like(b)
        info = P.SolveLinear(A, b, x, 2000, 1e-6, NORM2N, bicg, False, 0, False, False, False, False)[0]
        assert info.converged and info.isFiniteResidual
        out.append((x, info.usedIterations))
    assert out[0][1] == out[1][1] and torch.equal(out[0][0], out[1][0])
