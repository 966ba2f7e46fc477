"""pict_solve_linear_noalloc through the C ABI against pict_solve_linear: the same kernel on the same grid, so
every solution and iteration count is bitwise equal.  Two calls in a row on the scratch array, the second on a
smaller system (the first call's contents must not leak into it), for CG and BiCGStab with two right-hand sides;
a system whose workspace exceeds the scratch array is refused with PICT_ERR_INVALID before anything runs."""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from pict_amd import _abi

pytestmark = pytest.mark.gpu


def periodic_laplacian(shape, shift):
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


class System:
    """A CSR matrix and right-hand sides in device memory, and its C-ABI descriptor."""

    def __init__(self, m, nrhs, seed, dtype=torch.float32, nonsym=False):
        if nonsym:  # values no longer symmetric (BiCGStab systems)
            m = m.copy()
            m.data = m.data * (1 + 0.1 * np.random.default_rng(seed).random(m.data.shape))
        self.n = m.shape[0]
        self.val = torch.tensor(m.data, dtype=dtype).cuda()
        self.idx = torch.tensor(m.indices, dtype=torch.int32).cuda()
        self.row = torch.tensor(m.indptr, dtype=torch.int32).cuda()
        g = torch.Generator().manual_seed(seed)
        self.b = (torch.rand(nrhs * self.n, generator=g, dtype=torch.float64) * 2 - 1).to(dtype).cuda()
        self.nrhs = nrhs
        d = _abi.CsrDesc()
        d.n, d.nnz = self.n, m.nnz
        d.dtype = _abi.PICT_DTYPE_F32 if dtype == torch.float32 else _abi.PICT_DTYPE_F64
        d.row = ctypes.c_void_p(self.row.data_ptr())
        d.index = ctypes.c_void_p(self.idx.data_ptr())
        d.value = ctypes.c_void_p(self.val.data_ptr())
        d.transpose_perm = None
        self.desc = d


def opts(bicg, tol=1e-6, maxit=3000):
    o = _abi.SolveOpts()
    o.max_iterations = maxit
    o.convergence_criterion = 1  # NORM2_NORMALIZED
    o.use_bicg = int(bicg)
    o.tolerance = tol
    return o


def run(fn, sysm, o):
    x = torch.zeros_like(sysm.b)
    infos = (_abi.SolveInfo * sysm.nrhs)()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    st = fn(ctypes.byref(sysm.desc), ctypes.c_void_p(sysm.b.data_ptr()), ctypes.c_void_p(x.data_ptr()), sysm.nrhs,
            ctypes.byref(o), infos, stream)
    assert st == _abi.PICT_OK, _abi.load_library().pict_last_error().decode()
    return x, [(i.used_iterations, i.converged, i.final_residual) for i in infos]


@pytest.mark.parametrize("bicg", [False, True])
def test_noalloc_solve_matches_pict_solve_linear(bicg):
    lib = _abi.load_library()
    nrhs = 2 if bicg else 1
    big = System(periodic_laplacian((256, 192), 0.05), nrhs, 1, nonsym=bicg)
    small = System(periodic_laplacian((96, 64), 0.2), nrhs, 2, nonsym=bicg)
    o = opts(bicg)
    for case in (big, small):  # two calls in a row on the shared scratch, the second smaller
        print("case noalloc", "bicgstab" if bicg else "cg", case.n, "rows", nrhs, "rhs", flush=True)
        x_ref, i_ref = run(lib.pict_solve_linear, case, o)
        x_na, i_na = run(lib.pict_solve_linear_noalloc, case, o)
        torch.cuda.synchronize()
        assert all(c for _, c, _ in i_na), i_na
        assert i_na == i_ref
        assert torch.equal(x_na, x_ref)


def test_noalloc_refuses_systems_beyond_scratch():
    lib = _abi.load_library()
    sysm = System(periodic_laplacian((2048, 2048), 0.05), 1, 3)  # at least 36 bytes per row: 151 MB > 64 MiB
    x = torch.zeros_like(sysm.b)
    infos = (_abi.SolveInfo * 1)()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    st = lib.pict_solve_linear_noalloc(ctypes.byref(sysm.desc), ctypes.c_void_p(sysm.b.data_ptr()),
                                       ctypes.c_void_p(x.data_ptr()), 1, ctypes.byref(opts(False)), infos, stream)
    assert st == _abi.PICT_ERR_INVALID
    assert b"scratch" in lib.pict_last_error()
    assert float(x.abs().max()) == 0.0
