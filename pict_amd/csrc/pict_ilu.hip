// pict_ilu.hip -- the ILU(0)-preconditioned BiCGStab: SolveLinear with BiCGwithPreconditioner
// (bicgstabSolveGPU, bicgstab_solver_kernel.cu:139-210 for the factorization, :222-365 for the iteration), which the
// reference's drivers take as a retry when the unpreconditioned BiCGStab fails (PISOtorch_diff.py:276-286).
//
// The factorization and the triangular solves are sequential along dependency chains: a row of the lower factor
// waits for the rows it references below the diagonal, a row of the upper factor for those above.  The host turns
// the CSR pattern into level sets once per call (the analysis step); one workgroup then walks the levels with a
// workgroup barrier between them, first factoring, then running the whole iteration for each right-hand side.  A
// single workgroup needs no grid-wide synchronisation, and all of its waves share one L1, so a barrier orders every
// global write before the next level's reads.  It is a retry path: it is latency-bound (one barrier per level of
// every triangular solve) and uses one CU.
//
// The preconditioner is M = L U with L unit-lower and U upper on the matrix's own pattern, (L U)_ij = a_ij on that
// pattern, applied on the right: pw = M^-1 p, s = M^-1 r.  A transposed solve factors A^T, built on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "pict_host.hpp"

namespace pict {
namespace {

constexpr int kIluThreads = 1024;
constexpr int kIluWaves = kIluThreads / 64;

template <typename R>
struct IluArgs {
    int n, maxit, conv, nrhs, zero_shortcut, nl, nu;
    double tol;
    const int *row, *col, *diag;
    const int *lrows, *lptr, *urows, *uptr;  // rows by level of the lower / upper factor
    const R *val;
    R *lu;
    const R *b;
    R *x;
    R *r, *rw, *p, *pw, *v, *s, *t, *tmp;
    pict_solve_info *info;
    int *zero_all;
};

// Workgroup sums (or maxima, where is_max) of K values, returned to every thread.
template <int K>
__device__ __forceinline__ void wg_reduce(double (&v)[K], const bool (&is_max)[K]) {
    __shared__ double lds[kIluWaves][K];
#pragma unroll
    for (int k = 0; k < K; ++k)
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_xor(v[k], off);
            v[k] = is_max[k] ? fmax(v[k], o) : v[k] + o;
        }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) lds[wave][k] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double a = lds[0][k];
        for (int w = 1; w < kIluWaves; ++w) a = is_max[k] ? fmax(a, lds[w][k]) : a + lds[w][k];
        v[k] = a;
    }
    __syncthreads();
}

// ComputeConvergenceCriterion (cg_solver_kernel.cu:77-110) from sum r^2, sum |r|, max |r|
__device__ __forceinline__ double ilu_criterion(int conv, double rr, double asum, double amax, int n) {
    switch (conv) {
        case PICT_CONV_NORM2: return sqrt(rr);
        case PICT_CONV_NORM2_NORMALIZED: return sqrt(rr) / sqrt((double)n);
        case PICT_CONV_ABS_SUM: return asum;
        case PICT_CONV_ABS_MEAN: return asum / n;
        default: return amax;
    }
}

template <typename R>
__device__ __forceinline__ double residual_criterion(const IluArgs<R> &a) {
    double v[3] = {0, 0, 0};
    const bool mx[3] = {false, false, true};
    for (int i = threadIdx.x; i < a.n; i += kIluThreads) {
        const double ri = (double)a.r[i];
        v[0] += ri * ri;
        v[1] += fabs(ri);
        v[2] = fmax(v[2], fabs(ri));
    }
    wg_reduce<3>(v, mx);
    return ilu_criterion(a.conv, v[0], v[1], v[2], a.n);
}

// out = A in
template <typename R>
__device__ __forceinline__ void wg_spmv(const IluArgs<R> &a, const R *in, R *out) {
    __syncthreads();
    for (int i = threadIdx.x; i < a.n; i += kIluThreads) {
        R s = 0;
        for (int k = a.row[i]; k < a.row[i + 1]; ++k) s += a.val[k] * in[a.col[k]];
        out[i] = s;
    }
    __syncthreads();
}

// out = (L U)^-1 in: forward substitution with the unit-lower L into tmp, level by level, then back substitution
template <typename R>
__device__ __forceinline__ void wg_precondition(const IluArgs<R> &a, const R *in, R *out) {
    __syncthreads();
    for (int lev = 0; lev < a.nl; ++lev) {
        for (int q = a.lptr[lev] + threadIdx.x; q < a.lptr[lev + 1]; q += kIluThreads) {
            const int i = a.lrows[q];
            R s = in[i];
            for (int k = a.row[i]; k < a.diag[i]; ++k) s -= a.lu[k] * a.tmp[a.col[k]];
            a.tmp[i] = s;
        }
        __syncthreads();
    }
    for (int lev = 0; lev < a.nu; ++lev) {
        for (int q = a.uptr[lev] + threadIdx.x; q < a.uptr[lev + 1]; q += kIluThreads) {
            const int i = a.urows[q];
            R s = a.tmp[i];
            for (int k = a.diag[i] + 1; k < a.row[i + 1]; ++k) s -= a.lu[k] * out[a.col[k]];
            out[i] = s / a.lu[a.diag[i]];
        }
        __syncthreads();
    }
}

// ILU(0) in place on a copy of the values, level by level (IKJ per row: each earlier row k referenced below the
// diagonal scales l_ik = a_ik / u_kk and subtracts l_ik * u_kj from the row's later entries that row k also has);
// and the all-zero test of the right-hand sides for the shortcut.
template <typename R>
__global__ __launch_bounds__(kIluThreads) void k_ilu_factor(IluArgs<R> a) {
    const int nnz = a.row[a.n];
    for (int k = threadIdx.x; k < nnz; k += kIluThreads) a.lu[k] = a.val[k];
    double bmax[1] = {0};
    const bool mx[1] = {true};
    for (size_t i = threadIdx.x; i < (size_t)a.n * a.nrhs; i += kIluThreads) bmax[0] = fmax(bmax[0], fabs((double)a.b[i]));
    wg_reduce<1>(bmax, mx);  // its barriers also order the copy above before the factorization
    if (threadIdx.x == 0) *a.zero_all = (a.zero_shortcut && bmax[0] == 0.0) ? 1 : 0;
    for (int lev = 0; lev < a.nl; ++lev) {
        for (int q = a.lptr[lev] + threadIdx.x; q < a.lptr[lev + 1]; q += kIluThreads) {
            const int i = a.lrows[q], end = a.row[i + 1];
            for (int kk = a.row[i]; kk < a.diag[i]; ++kk) {
                const int k = a.col[kk];
                const R l = a.lu[kk] / a.lu[a.diag[k]];
                a.lu[kk] = l;
                int qk = a.diag[k] + 1;
                const int kend = a.row[k + 1];
                for (int jj = kk + 1; jj < end; ++jj) {
                    const int j = a.col[jj];
                    while (qk < kend && a.col[qk] < j) ++qk;
                    if (qk < kend && a.col[qk] == j) a.lu[jj] -= l * a.lu[qk];
                }
            }
        }
        __syncthreads();
    }
}

// The right-preconditioned BiCGStab of one right-hand side (bicgstab_solver_kernel.cu:240-345): the stopping rule
// and the reported iterations of the unpreconditioned solver, dot products in FP64.
template <typename R>
__global__ __launch_bounds__(kIluThreads) void k_bicgstab_ilu(IluArgs<R> a, int rhs) {
    const int n = a.n;
    const R *b = a.b + (size_t)rhs * n;
    R *x = a.x + (size_t)rhs * n;
    pict_solve_info *info = a.info + rhs;
    if (*a.zero_all) {
        for (int i = threadIdx.x; i < n; i += kIluThreads) x[i] = 0;
        if (threadIdx.x == 0) {
            info->final_residual = 0;
            info->used_iterations = -1;
            info->converged = 1;
            info->is_finite_residual = 1;
        }
        return;
    }
    wg_spmv(a, x, a.r);
    for (int i = threadIdx.x; i < n; i += kIluThreads) {
        const R ri = b[i] - a.r[i];
        a.r[i] = ri;
        a.rw[i] = ri;
        a.p[i] = ri;
        a.v[i] = 0;
    }
    double final_res = 0;
    int used = -1, converged = 0, finite = 1;
    const double nrmr0 = residual_criterion(a);
    if (!(nrmr0 < a.tol)) {
        double rho = 0, alpha = 0, omega = 0;
        const bool sum1[1] = {false};
        const bool sum2[2] = {false, false};
        for (int it = 0; it < a.maxit; ++it) {
            const double rhop = rho;
            double d1[1] = {0};
            for (int i = threadIdx.x; i < n; i += kIluThreads) d1[0] += (double)a.rw[i] * (double)a.r[i];
            wg_reduce<1>(d1, sum1);
            rho = d1[0];
            if (it > 0) {
                const R beta = (R)((rho / rhop) * (alpha / omega)), om = (R)omega;
                for (int i = threadIdx.x; i < n; i += kIluThreads) a.p[i] = a.r[i] + beta * (a.p[i] - om * a.v[i]);
            }
            wg_precondition(a, a.p, a.pw);
            wg_spmv(a, a.pw, a.v);
            d1[0] = 0;
            for (int i = threadIdx.x; i < n; i += kIluThreads) d1[0] += (double)a.rw[i] * (double)a.v[i];
            wg_reduce<1>(d1, sum1);
            alpha = rho / d1[0];
            {
                const R al = (R)alpha;
                for (int i = threadIdx.x; i < n; i += kIluThreads) {
                    a.r[i] -= al * a.v[i];
                    x[i] += al * a.pw[i];
                }
            }
            double nr = residual_criterion(a);
            if (!isfinite(nr)) {
                converged = 0;
                used = it;
                final_res = nr;
                finite = 0;
                break;
            }
            used = it;
            final_res = nr;
            if (nr < a.tol) {
                converged = 1;
                break;
            }
            wg_precondition(a, a.r, a.s);
            wg_spmv(a, a.s, a.t);
            double d2[2] = {0, 0};
            for (int i = threadIdx.x; i < n; i += kIluThreads) {
                d2[0] += (double)a.t[i] * (double)a.r[i];
                d2[1] += (double)a.t[i] * (double)a.t[i];
            }
            wg_reduce<2>(d2, sum2);
            omega = d2[0] / d2[1];
            {
                const R om = (R)omega;
                for (int i = threadIdx.x; i < n; i += kIluThreads) {
                    x[i] += om * a.s[i];
                    a.r[i] -= om * a.t[i];
                }
            }
            nr = residual_criterion(a);
            final_res = nr;
            if (nr < a.tol) {
                converged = 1;
                used = it + 1;
                break;
            }
        }
    } else {
        converged = 1;
        final_res = nrmr0;
    }
    if (threadIdx.x == 0) {
        info->final_residual = final_res;
        info->used_iterations = used;
        info->converged = converged;
        info->is_finite_residual = finite;
    }
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// rows of each level: ptr[l] .. ptr[l+1] in rows
void bucket_levels(const std::vector<int> &level, int nlev, std::vector<int> &rows, std::vector<int> &ptr) {
    ptr.assign((size_t)nlev + 1, 0);
    for (int l : level) ptr[(size_t)l + 1]++;
    for (int l = 0; l < nlev; ++l) ptr[(size_t)l + 1] += ptr[l];
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    rows.resize(level.size());
    for (size_t i = 0; i < level.size(); ++i) rows[(size_t)fill[level[i]]++] = (int)i;
}

template <typename R>
pict_status ilu_solve_typed(const pict_csr_desc *A, const void *rhs, void *x, int nrhs, const pict_solve_opts *o,
                            pict_solve_info *infos, hipStream_t s) {
    const int n = A->n;
    std::vector<int> hrow((size_t)n + 1);
    PICT_HIP_TRY(hipMemcpyAsync(hrow.data(), A->row, sizeof(int) * ((size_t)n + 1), hipMemcpyDeviceToHost, s));
    PICT_HIP_TRY(hipStreamSynchronize(s));
    const int nnz = hrow[n];
    PICT_CHECK_ARG(nnz >= n && hrow[0] == 0, "ILU(0): invalid CSR row pointers");
    for (int i = 0; i < n; ++i)
        PICT_CHECK_ARG(hrow[i] <= hrow[(size_t)i + 1] && hrow[(size_t)i + 1] <= nnz,
                       "ILU(0): CSR row pointers must be non-decreasing (row %d)", i);
    std::vector<int> hcol((size_t)nnz);
    std::vector<R> hval((size_t)nnz);
    PICT_HIP_TRY(hipMemcpyAsync(hcol.data(), A->index, sizeof(int) * (size_t)nnz, hipMemcpyDeviceToHost, s));
    PICT_HIP_TRY(hipMemcpyAsync(hval.data(), A->value, sizeof(R) * (size_t)nnz, hipMemcpyDeviceToHost, s));
    PICT_HIP_TRY(hipStreamSynchronize(s));
    // the matrix the iteration works on: A, or A^T with ascending columns per row
    std::vector<int> brow, bcol;
    std::vector<R> bval;
    if (!o->transpose_a) {
        brow.swap(hrow);
        bcol.swap(hcol);
        bval.swap(hval);
    } else {
        brow.assign((size_t)n + 1, 0);
        for (int k = 0; k < nnz; ++k) {
            PICT_CHECK_ARG(hcol[k] >= 0 && hcol[k] < n, "ILU(0): column index out of range");
            brow[(size_t)hcol[k] + 1]++;
        }
        for (int i = 0; i < n; ++i) brow[(size_t)i + 1] += brow[i];
        std::vector<int> fill(brow.begin(), brow.end() - 1);
        bcol.resize((size_t)nnz);
        bval.resize((size_t)nnz);
        for (int i = 0; i < n; ++i)
            for (int k = hrow[i]; k < hrow[(size_t)i + 1]; ++k) {
                const int d = fill[hcol[k]]++;
                bcol[d] = i;
                bval[d] = hval[k];
            }
    }
    // analysis: ascending columns, the diagonal's position, levels of both factors
    std::vector<int> diag((size_t)n), lf((size_t)n), lb((size_t)n);
    int nl = 0, nu = 0;
    for (int i = 0; i < n; ++i) {
        int prev = -1, d = -1;
        for (int k = brow[i]; k < brow[(size_t)i + 1]; ++k) {
            const int c = bcol[k];
            PICT_CHECK_ARG(c > prev && c < n, "ILU(0): row %d: column indices must be ascending and in range", i);
            prev = c;
            if (c == i) d = k;
        }
        PICT_CHECK_ARG(d >= 0, "ILU(0): A(%d,%d) is missing", i, i);
        diag[i] = d;
        int l = 0;
        for (int k = brow[i]; k < d; ++k) l = std::max(l, lf[bcol[k]] + 1);
        lf[i] = l;
        nl = std::max(nl, l + 1);
    }
    for (int i = n - 1; i >= 0; --i) {
        int l = 0;
        for (int k = diag[i] + 1; k < brow[(size_t)i + 1]; ++k) l = std::max(l, lb[bcol[k]] + 1);
        lb[i] = l;
        nu = std::max(nu, l + 1);
    }
    std::vector<int> lrows, lptr, urows, uptr;
    bucket_levels(lf, nl, lrows, lptr);
    bucket_levels(lb, nu, urows, uptr);

    // device workspace
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off += align256(bytes);
        return at;
    };
    const size_t o_row = take(sizeof(int) * ((size_t)n + 1)), o_col = take(sizeof(int) * (size_t)nnz),
                 o_diag = take(sizeof(int) * (size_t)n), o_lrows = take(sizeof(int) * (size_t)n),
                 o_lptr = take(sizeof(int) * ((size_t)nl + 1)), o_urows = take(sizeof(int) * (size_t)n),
                 o_uptr = take(sizeof(int) * ((size_t)nu + 1)), o_flag = take(sizeof(int)),
                 o_info = take(sizeof(pict_solve_info) * (size_t)nrhs), o_val = take(sizeof(R) * (size_t)nnz),
                 o_lu = take(sizeof(R) * (size_t)nnz), o_vec = take(sizeof(R) * (size_t)n * 8);
    char *ws = nullptr;
    PICT_HIP_TRY(ws_alloc((void **)&ws, off, s));
    struct Guard {
        char *p;
        hipStream_t s;
        ~Guard() {
            if (p) {
                (void)hipStreamSynchronize(s);
                (void)ws_free(p, s);
            }
        }
    } guard{ws, s};
    PICT_HIP_TRY(hipMemcpyAsync(ws + o_row, brow.data(), sizeof(int) * ((size_t)n + 1), hipMemcpyHostToDevice, s));
    PICT_HIP_TRY(hipMemcpyAsync(ws + o_col, bcol.data(), sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice, s));
    PICT_HIP_TRY(hipMemcpyAsync(ws + o_diag, diag.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, s));
    PICT_HIP_TRY(hipMemcpyAsync(ws + o_lrows, lrows.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, s));
    PICT_HIP_TRY(hipMemcpyAsync(ws + o_lptr, lptr.data(), sizeof(int) * ((size_t)nl + 1), hipMemcpyHostToDevice, s));
    PICT_HIP_TRY(hipMemcpyAsync(ws + o_urows, urows.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, s));
    PICT_HIP_TRY(hipMemcpyAsync(ws + o_uptr, uptr.data(), sizeof(int) * ((size_t)nu + 1), hipMemcpyHostToDevice, s));
    PICT_HIP_TRY(hipMemcpyAsync(ws + o_val, bval.data(), sizeof(R) * (size_t)nnz, hipMemcpyHostToDevice, s));

    IluArgs<R> a;
    a.n = n;
    a.maxit = o->max_iterations;
    a.conv = o->convergence_criterion;
    a.nrhs = nrhs;
    a.zero_shortcut = o->zero_rhs_shortcut;
    a.nl = nl;
    a.nu = nu;
    a.tol = o->tolerance;
    a.row = (const int *)(ws + o_row);
    a.col = (const int *)(ws + o_col);
    a.diag = (const int *)(ws + o_diag);
    a.lrows = (const int *)(ws + o_lrows);
    a.lptr = (const int *)(ws + o_lptr);
    a.urows = (const int *)(ws + o_urows);
    a.uptr = (const int *)(ws + o_uptr);
    a.val = (const R *)(ws + o_val);
    a.lu = (R *)(ws + o_lu);
    a.b = (const R *)rhs;
    a.x = (R *)x;
    R *vec = (R *)(ws + o_vec);
    a.r = vec;
    a.rw = vec + n;
    a.p = vec + 2 * (size_t)n;
    a.pw = vec + 3 * (size_t)n;
    a.v = vec + 4 * (size_t)n;
    a.s = vec + 5 * (size_t)n;
    a.t = vec + 6 * (size_t)n;
    a.tmp = vec + 7 * (size_t)n;
    a.info = (pict_solve_info *)(ws + o_info);
    a.zero_all = (int *)(ws + o_flag);
    hipLaunchKernelGGL(k_ilu_factor<R>, dim3(1), dim3(kIluThreads), 0, s, a);
    PICT_HIP_TRY(hipGetLastError());
    for (int k = 0; k < nrhs; ++k) {
        hipLaunchKernelGGL(k_bicgstab_ilu<R>, dim3(1), dim3(kIluThreads), 0, s, a, k);
        PICT_HIP_TRY(hipGetLastError());
    }
    PICT_HIP_TRY(hipMemcpyAsync(infos, ws + o_info, sizeof(pict_solve_info) * (size_t)nrhs, hipMemcpyDeviceToHost, s));
    PICT_HIP_TRY(hipStreamSynchronize(s));
    return PICT_OK;
}

}  // namespace

pict_status ilu_bicgstab_solve(const pict_csr_desc *A, const void *rhs, void *x, int nrhs, const pict_solve_opts *o,
                               pict_solve_info *infos, hipStream_t s) {
    if (A->dtype == PICT_DTYPE_F32) return ilu_solve_typed<float>(A, rhs, x, nrhs, o, infos, s);
    return ilu_solve_typed<double>(A, rhs, x, nrhs, o, infos, s);
}

}  // namespace pict
