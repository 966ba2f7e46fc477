"""The ILU(0)-preconditioned BiCGStab in the oracle (bicgstabSolveGPU with preconditioner,
bicgstab_solver_kernel.cu:139-210 for the factorization, :222-365 for the iteration), pinned by what the
mathematics fixes:
- the ILU(0) identity: with L unit-lower and U upper on the matrix's pattern, (L U)_ij = a_ij for every (i,j) of the
  pattern; checked on the advection and pressure matrices of every scenario, and on A^T for transposed solves (a
  dropped pivot division or a skipped update fails it);
- a tridiagonal matrix has no fill, so ILU(0) is the exact LU factorization and the preconditioned iteration stops
  at its first half step;
- the returned x satisfies the stopping criterion on b - A x (and b - A^T x), agrees with the unpreconditioned
  solution to the tolerance, and takes fewer iterations on a pressure matrix;
- a missing diagonal entry is an error, and an all-zero right-hand side gives x = 0 with no iterations.
The reference's factor descriptors set the upper factor's fill mode and diagonal type on the lower factor's
descriptor (:157-163), so which triangular systems its cuSPARSE path solves is not fixed by its source; the
restatement applies the preconditioner those lines declare, M = L U.
"""
import numpy as np
import pytest
import torch

import scenarios
from oracle import oracle as O
from pict_amd import PISOtorch as P


def _matrices(name):
    d = scenarios.ALL[name](dtype=torch.float64, device="cpu") if name in scenarios.ALL else \
        scenarios.NON_ORTHO[name](dtype=torch.float64, device="cpu")
    O.build_csr_pattern(d)
    O.SetupAdvectionMatrix(d, 0.05)
    O.SetupPressureCorrection(d, 0.05)
    return d


def _dense(A, vals=None):
    n = A.getRows()
    row, col = A.row.numpy(), A.index.numpy()
    v = (A.value if vals is None else vals).numpy()
    M = np.zeros((n, n))
    for i in range(n):
        M[i, col[row[i]:row[i + 1]]] = v[row[i]:row[i + 1]]
    return M


def _transpose_pattern(A):
    """CSR of A^T (rows ascending in columns), as a CSRmatrix"""
    n = A.getRows()
    M = _dense(A).T
    rows, cols, vals = [0], [], []
    pat = _dense(P.CSRmatrix(torch.ones_like(A.value), A.index, A.row)).T != 0
    for i in range(n):
        js = np.nonzero(pat[i])[0]
        cols += list(js)
        vals += list(M[i, js])
        rows.append(len(cols))
    return P.CSRmatrix(torch.tensor(vals, dtype=torch.float64), torch.tensor(cols, dtype=torch.int32),
                       torch.tensor(rows, dtype=torch.int32))


def _check_ilu_identity(A, lu, pattern_of):
    n = A.getRows()
    F = _dense(pattern_of, lu)
    Lm = np.tril(F, -1) + np.eye(n)
    Um = np.triu(F)
    prod = Lm @ Um
    Ad = _dense(pattern_of)
    mask = _dense(P.CSRmatrix(torch.ones_like(pattern_of.value), pattern_of.index, pattern_of.row)) != 0
    np.testing.assert_allclose(prod[mask], Ad[mask], rtol=0, atol=1e-12 * np.abs(Ad).max())


CASES = ["cavity", "periodic_two_block", "rotated_connection", "box3d", "three_d_connected", "sheared_walls",
         "sheared_two_block_rotated"]


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("which", ["C", "P"])
def test_ilu0_reproduces_the_matrix_on_its_pattern(name, which):
    d = _matrices(name)
    A = d.C if which == "C" else d.P
    _check_ilu_identity(A, O.ILU0(A), A)
    AT = _transpose_pattern(A)
    _check_ilu_identity(AT, O.ILU0(A, transpose=True), AT)


def _tridiagonal(n, seed):
    g = torch.Generator().manual_seed(seed)
    rows, cols, vals = [0], [], []
    for i in range(n):
        for j in (i - 1, i, i + 1):
            if 0 <= j < n:
                cols.append(j)
                vals.append(4.0 + float(torch.rand(1, generator=g)) if i == j else -1.0 - float(torch.rand(1, generator=g)))
        rows.append(len(cols))
    return P.CSRmatrix(torch.tensor(vals, dtype=torch.float64), torch.tensor(cols, dtype=torch.int32),
                       torch.tensor(rows, dtype=torch.int32))


def test_tridiagonal_ilu_is_exact_and_solves_in_one_half_step():
    A = _tridiagonal(40, 3)
    lu = O.ILU0(A)
    n = A.getRows()
    F = _dense(A, lu)
    np.testing.assert_allclose((np.tril(F, -1) + np.eye(n)) @ np.triu(F), _dense(A), rtol=0, atol=1e-13)
    b = torch.randn(n, generator=torch.Generator().manual_seed(4), dtype=torch.float64)
    x = torch.zeros(n, dtype=torch.float64)
    info = O.SolveLinear(A, b, x, 100, 1e-10, 1, useBiCG=True, withPreconditioner=True)[0]
    assert info.converged and info.usedIterations == 0
    np.testing.assert_allclose(_dense(A) @ x.numpy(), b.numpy(), rtol=0, atol=1e-12)


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("transpose", [False, True])
def test_preconditioned_solution_meets_the_criterion(name, transpose):
    d = _matrices(name)
    A = d.C
    n = A.getRows()
    nrhs = d.getSpatialDims()
    b = torch.randn(n * nrhs, generator=torch.Generator().manual_seed(5), dtype=torch.float64)
    x = torch.zeros_like(b)
    tol = 1e-10
    infos = O.SolveLinear(A, b, x, 2000, tol, 1, useBiCG=True, transposeA=transpose, withPreconditioner=True)
    M = _dense(A).T if transpose else _dense(A)
    y = torch.zeros_like(b)
    plain = O.SolveLinear(A, b, y, 2000, tol, 1, useBiCG=True, transposeA=transpose)
    for k, info in enumerate(infos):
        assert info.converged, info
        xs = x.numpy()[k * n:(k + 1) * n]
        r = b.numpy()[k * n:(k + 1) * n] - M @ xs
        assert np.linalg.norm(r) / np.sqrt(n) < tol
        assert np.abs(xs - y.numpy()[k * n:(k + 1) * n]).max() < 1e-7 * max(1.0, np.abs(xs).max())
        assert info.usedIterations <= plain[k].usedIterations


def test_preconditioner_cuts_iterations_on_a_pressure_matrix():
    d = _matrices("box3d")
    A = d.P
    n = A.getRows()
    b = torch.randn(n, generator=torch.Generator().manual_seed(6), dtype=torch.float64)
    b -= b.mean()  # the pressure matrix of a closed box is singular; a compatible right-hand side
    x, y = torch.zeros_like(b), torch.zeros_like(b)
    pre = O.SolveLinear(A, b, x, 2000, 1e-9, 1, useBiCG=True, withPreconditioner=True)[0]
    plain = O.SolveLinear(A, b, y, 2000, 1e-9, 1, useBiCG=True)[0]
    assert pre.converged and plain.converged
    assert pre.usedIterations <= 0.7 * plain.usedIterations, (pre.usedIterations, plain.usedIterations)


def test_missing_diagonal_is_an_error():
    A = P.CSRmatrix(torch.tensor([1.0, 2.0, 3.0], dtype=torch.float64), torch.tensor([1, 0, 1], dtype=torch.int32),
                    torch.tensor([0, 1, 3], dtype=torch.int32))
    with pytest.raises(RuntimeError):
        O.ILU0(A)
    with pytest.raises(RuntimeError):
        O.SolveLinear(A, torch.ones(2, dtype=torch.float64), torch.zeros(2, dtype=torch.float64), 10, 1e-8, 1,
                      useBiCG=True, withPreconditioner=True)
