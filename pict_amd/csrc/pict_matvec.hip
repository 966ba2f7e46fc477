// pict_matvec.hip -- per-cell products of vectors and small square matrices held in the
// channel dimension of NCDHW tensors, and their vector-Jacobian products
// (matrix_vector_ops.cu:7-303, matrix_vector_ops_grads.cu:8-310).
//
// A tensor [1, C, (Z), Y, X] holds per cell a vector (C = D) or a row-major D x D matrix
// (C = D*D); channel c of cell i is at c*cells + i.  One thread per cell: every channel
// plane is read and written with consecutive threads on consecutive cells.  The work is
// a few flops per loaded value, so the kernels are bound by HBM bandwidth.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "pict_host.hpp"

namespace pict {

template <typename R, int D>
struct SqMat {
    R m[D][D];
};

template <typename R, int D>
__device__ __forceinline__ SqMat<R, D> load_mat(const R *p, size_t i, size_t cells, bool transpose) {
    SqMat<R, D> a;
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) {
            const R v = p[(size_t)(r * D + c) * cells + i];
            if (transpose) a.m[c][r] = v;
            else a.m[r][c] = v;
        }
    return a;
}
template <typename R, int D>
__device__ __forceinline__ void store_mat(R *p, size_t i, size_t cells, const SqMat<R, D> &a, bool transpose) {
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) p[(size_t)(r * D + c) * cells + i] = transpose ? a.m[c][r] : a.m[r][c];
}
template <typename R, int D>
__device__ __forceinline__ void load_vec(const R *p, size_t i, size_t cells, R (&v)[D]) {
#pragma unroll
    for (int k = 0; k < D; ++k) v[k] = p[(size_t)k * cells + i];
}
template <typename R, int D>
__device__ __forceinline__ void store_vec(R *p, size_t i, size_t cells, const R (&v)[D]) {
#pragma unroll
    for (int k = 0; k < D; ++k) p[(size_t)k * cells + i] = v[k];
}

// inverse = adjugate / determinant (transformations.h:418-482: the cofactor expansion along the first row)
template <typename R>
__device__ __forceinline__ SqMat<R, 1> inverse(const SqMat<R, 1> &a) {
    SqMat<R, 1> b;
    b.m[0][0] = R(1) * (R(1) / a.m[0][0]);
    return b;
}
template <typename R>
__device__ __forceinline__ SqMat<R, 2> inverse(const SqMat<R, 2> &a) {
    const R s = R(1) / (a.m[0][0] * a.m[1][1] - a.m[1][0] * a.m[0][1]);
    SqMat<R, 2> b;
    b.m[0][0] = a.m[1][1] * s;
    b.m[0][1] = -a.m[0][1] * s;
    b.m[1][0] = -a.m[1][0] * s;
    b.m[1][1] = a.m[0][0] * s;
    return b;
}
template <typename R>
__device__ __forceinline__ SqMat<R, 3> inverse(const SqMat<R, 3> &a) {
    const auto &m = a.m;
    // cofactors of the first row give the determinant; the adjugate is the transposed cofactor matrix
    const R c00 = m[1][1] * m[2][2] - m[2][1] * m[1][2];
    const R c01 = m[1][0] * m[2][2] - m[2][0] * m[1][2];
    const R c02 = m[1][0] * m[2][1] - m[2][0] * m[1][1];
    const R s = R(1) / (m[0][0] * c00 - m[0][1] * c01 + m[0][2] * c02);
    SqMat<R, 3> b;
    b.m[0][0] = c00 * s;
    b.m[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * s;
    b.m[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * s;
    b.m[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * s;
    b.m[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * s;
    b.m[1][2] = (m[1][0] * m[0][2] - m[0][0] * m[1][2]) * s;
    b.m[2][0] = c02 * s;
    b.m[2][1] = (m[2][0] * m[0][1] - m[0][0] * m[2][1]) * s;
    b.m[2][2] = (m[0][0] * m[1][1] - m[1][0] * m[0][1]) * s;
    return b;
}

// operand as the product sees it: loaded (transposed), then inverted
template <typename R, int D>
__device__ __forceinline__ SqMat<R, D> operand(const R *p, size_t i, size_t cells, bool transpose, bool invert) {
    SqMat<R, D> a = load_mat<R, D>(p, i, cells, transpose);
    return invert ? inverse(a) : a;
}

enum MatmulKind { kVV = 0, kMV = 1, kVM = 2, kMM = 3 };

// out = A B per cell.  Matrix x matrix honours every flag; matrix x vector and vector x matrix the flags
// of their matrix only; vector x vector none (k_mul_*, matrix_vector_ops.cu:7-111)
template <typename R, int D, int KIND>
__global__ __launch_bounds__(256) void k_matmul(const R *__restrict__ A, const R *__restrict__ B, R *__restrict__ out,
                                                size_t cells, int flags) {
    const bool tA = flags & PICT_MATMUL_TRANSPOSE_A, iA = flags & PICT_MATMUL_INVERT_A;
    const bool tB = flags & PICT_MATMUL_TRANSPOSE_B, iB = flags & PICT_MATMUL_INVERT_B;
    const bool tO = flags & PICT_MATMUL_TRANSPOSE_OUT, iO = flags & PICT_MATMUL_INVERT_OUT;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < cells; i += (size_t)gridDim.x * blockDim.x) {
        if constexpr (KIND == kMM) {
            const SqMat<R, D> a = operand<R, D>(A, i, cells, tA, iA), b = operand<R, D>(B, i, cells, tB, iB);
            SqMat<R, D> c;
#pragma unroll
            for (int r = 0; r < D; ++r)
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    R s = 0;
#pragma unroll
                    for (int j = 0; j < D; ++j) s += b.m[j][k] * a.m[r][j];
                    c.m[r][k] = s;
                }
            if (iO) c = inverse(c);
            store_mat<R, D>(out, i, cells, c, tO);
        } else if constexpr (KIND == kMV) {
            const SqMat<R, D> a = operand<R, D>(A, i, cells, tA, iA);
            R v[D], o[D];
            load_vec<R, D>(B, i, cells, v);
#pragma unroll
            for (int r = 0; r < D; ++r) {
                R s = 0;
#pragma unroll
                for (int j = 0; j < D; ++j) s += a.m[r][j] * v[j];
                o[r] = s;
            }
            store_vec<R, D>(out, i, cells, o);
        } else if constexpr (KIND == kVM) {
            const SqMat<R, D> b = operand<R, D>(B, i, cells, tB, iB);
            R v[D], o[D];
            load_vec<R, D>(A, i, cells, v);
#pragma unroll
            for (int k = 0; k < D; ++k) {
                R s = 0;
#pragma unroll
                for (int j = 0; j < D; ++j) s += b.m[j][k] * v[j];
                o[k] = s;
            }
            store_vec<R, D>(out, i, cells, o);
        } else {
            R u[D], v[D];
            load_vec<R, D>(A, i, cells, u);
            load_vec<R, D>(B, i, cells, v);
            R s = 0;
#pragma unroll
            for (int j = 0; j < D; ++j) s += u[j] * v[j];
            out[i] = s;
        }
    }
}

// VJP of k_matmul for output gradient G (k_mul_*_GRAD, matrix_vector_ops_grads.cu:8-207).  An inverted
// operand gets no gradient (zeros); matrix x matrix gives none at all with an inverted output or with both
// operands inverted.  A transposed operand's gradient is written transposed (the gradient of the stored data).
template <typename R, int D, int KIND>
__global__ __launch_bounds__(256) void k_matmul_grad(const R *__restrict__ A, const R *__restrict__ B,
                                                     const R *__restrict__ G, R *__restrict__ gA, R *__restrict__ gB,
                                                     size_t cells, int flags) {
    const bool tA = flags & PICT_MATMUL_TRANSPOSE_A, iA = flags & PICT_MATMUL_INVERT_A;
    const bool tB = flags & PICT_MATMUL_TRANSPOSE_B, iB = flags & PICT_MATMUL_INVERT_B;
    const bool tO = flags & PICT_MATMUL_TRANSPOSE_OUT, iO = flags & PICT_MATMUL_INVERT_OUT;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < cells; i += (size_t)gridDim.x * blockDim.x) {
        if constexpr (KIND == kMM) {
            SqMat<R, D> da{}, db{};
            if (!(iO || (iA && iB))) {
                const SqMat<R, D> g = load_mat<R, D>(G, i, cells, tO);
                if (!iA) {  // dA = G B^T
                    const SqMat<R, D> b = operand<R, D>(B, i, cells, tB, iB);
#pragma unroll
                    for (int r = 0; r < D; ++r)
#pragma unroll
                        for (int c = 0; c < D; ++c) {
                            R s = 0;
#pragma unroll
                            for (int k = 0; k < D; ++k) s += b.m[c][k] * g.m[r][k];
                            da.m[r][c] = s;
                        }
                }
                if (!iB) {  // dB = A^T G
                    const SqMat<R, D> a = operand<R, D>(A, i, cells, tA, iA);
#pragma unroll
                    for (int r = 0; r < D; ++r)
#pragma unroll
                        for (int c = 0; c < D; ++c) {
                            R s = 0;
#pragma unroll
                            for (int k = 0; k < D; ++k) s += a.m[k][r] * g.m[k][c];
                            db.m[r][c] = s;
                        }
                }
            }
            store_mat<R, D>(gA, i, cells, da, tA);
            store_mat<R, D>(gB, i, cells, db, tB);
        } else if constexpr (KIND == kMV) {
            R g[D], v[D], dv[D];
            load_vec<R, D>(G, i, cells, g);
            load_vec<R, D>(B, i, cells, v);
            SqMat<R, D> da{};
            if (!iA)  // dA = g v^T
#pragma unroll
                for (int r = 0; r < D; ++r)
#pragma unroll
                    for (int c = 0; c < D; ++c) da.m[r][c] = v[c] * g[r];
            store_mat<R, D>(gA, i, cells, da, tA);
            const SqMat<R, D> a = operand<R, D>(A, i, cells, tA, iA);  // dv = A^T g
#pragma unroll
            for (int c = 0; c < D; ++c) {
                R s = 0;
#pragma unroll
                for (int k = 0; k < D; ++k) s += a.m[k][c] * g[k];
                dv[c] = s;
            }
            store_vec<R, D>(gB, i, cells, dv);
        } else if constexpr (KIND == kVM) {
            R g[D], u[D], du[D];
            load_vec<R, D>(G, i, cells, g);
            load_vec<R, D>(A, i, cells, u);
            const SqMat<R, D> b = operand<R, D>(B, i, cells, tB, iB);  // du = B g
#pragma unroll
            for (int r = 0; r < D; ++r) {
                R s = 0;
#pragma unroll
                for (int k = 0; k < D; ++k) s += b.m[r][k] * g[k];
                du[r] = s;
            }
            store_vec<R, D>(gA, i, cells, du);
            SqMat<R, D> db{};
            if (!iB)  // dB = u g^T
#pragma unroll
                for (int r = 0; r < D; ++r)
#pragma unroll
                    for (int c = 0; c < D; ++c) db.m[r][c] = u[r] * g[c];
            store_mat<R, D>(gB, i, cells, db, tB);
        } else {
            R u[D], v[D];
            load_vec<R, D>(A, i, cells, u);
            load_vec<R, D>(B, i, cells, v);
            const R g = G[i];
#pragma unroll
            for (int j = 0; j < D; ++j) {
                u[j] *= g;
                v[j] *= g;
            }
            store_vec<R, D>(gA, i, cells, v);  // d(u.v)/du = v
            store_vec<R, D>(gB, i, cells, u);
        }
    }
}

// matrices with the vector on the diagonal, zeros elsewhere (k_vectorToDiagMatrix, matrix_vector_ops.cu:203-232)
template <typename R, int D>
__global__ __launch_bounds__(256) void k_vec_to_diag(const R *__restrict__ v, R *__restrict__ out, size_t cells) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < cells; i += (size_t)gridDim.x * blockDim.x) {
        R d[D];
        load_vec<R, D>(v, i, cells, d);
        SqMat<R, D> m{};
#pragma unroll
        for (int k = 0; k < D; ++k) m.m[k][k] = d[k];
        store_mat<R, D>(out, i, cells, m, false);
    }
}

// per-cell inverse (k_inverseMatrix, matrix_vector_ops.cu:265-279); in and out may be the same tensor: a
// thread reads its cell's channels before it writes them
template <typename R, int D>
__global__ __launch_bounds__(256) void k_invert(const R *in, R *out, size_t cells) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < cells; i += (size_t)gridDim.x * blockDim.x)
        store_mat<R, D>(out, i, cells, inverse(load_mat<R, D>(in, i, cells, false)), false);
}

}  // namespace pict

using namespace pict;

namespace {

unsigned grid_for(size_t cells) {
    const size_t b = (cells + 255) / 256;
    return (unsigned)(b < 8192 ? (b ? b : 1) : 8192);
}

pict_status check_common(int32_t num_dims, int64_t cells, int32_t dtype) {
    PICT_CHECK_ARG(num_dims >= 1 && num_dims <= 3, "num_dims must be 1..3 (got %d)", num_dims);
    PICT_CHECK_ARG(cells >= 0, "cells must be >= 0");
    PICT_CHECK_ARG(dtype == PICT_DTYPE_F32 || dtype == PICT_DTYPE_F64, "unknown dtype %d", dtype);
    return PICT_OK;
}

// dtype x dimension x kind dispatch of a matmul kernel
#define PICT_MATMUL_DISPATCH(dtype, D_, kind, CALL)                                                  \
    do {                                                                                             \
        auto _go = [&](auto r_tag, auto d_tag, auto k_tag) {                                         \
            using R = decltype(r_tag);                                                               \
            constexpr int D = decltype(d_tag)::value;                                                \
            constexpr int K = decltype(k_tag)::value;                                                \
            CALL;                                                                                    \
        };                                                                                           \
        auto _k = [&](auto r_tag, auto d_tag) {                                                      \
            switch (kind) {                                                                          \
                case kVV: _go(r_tag, d_tag, std::integral_constant<int, kVV>{}); break;              \
                case kMV: _go(r_tag, d_tag, std::integral_constant<int, kMV>{}); break;              \
                case kVM: _go(r_tag, d_tag, std::integral_constant<int, kVM>{}); break;              \
                default: _go(r_tag, d_tag, std::integral_constant<int, kMM>{}); break;               \
            }                                                                                        \
        };                                                                                           \
        auto _d = [&](auto r_tag) {                                                                  \
            if ((D_) == 1) _k(r_tag, std::integral_constant<int, 1>{});                              \
            else if ((D_) == 2) _k(r_tag, std::integral_constant<int, 2>{});                         \
            else _k(r_tag, std::integral_constant<int, 3>{});                                        \
        };                                                                                           \
        if ((dtype) == PICT_DTYPE_F32) _d(float{});                                                  \
        else _d(double{});                                                                           \
    } while (0)

}  // namespace

extern "C" {

pict_status pict_matmul(const void *a, int32_t a_is_matrix, const void *b, int32_t b_is_matrix, void *out,
                        int32_t num_dims, int64_t cells, int32_t dtype, int32_t flags, void *stream) {
    pict_status st = check_common(num_dims, cells, dtype);
    if (st) return st;
    PICT_CHECK_ARG(a && b && out, "NULL tensor");
    if (cells == 0) return PICT_OK;
    const int kind = (a_is_matrix ? 1 : 0) + (b_is_matrix ? 2 : 0);
    hipStream_t s = (hipStream_t)stream;
    const unsigned g = grid_for((size_t)cells);
    PICT_MATMUL_DISPATCH(dtype, num_dims, kind,
                         hipLaunchKernelGGL((k_matmul<R, D, K>), dim3(g), dim3(256), 0, s, (const R *)a, (const R *)b,
                                            (R *)out, (size_t)cells, (int)flags));
    return check_launch();
}

pict_status pict_matmul_grad(const void *a, int32_t a_is_matrix, const void *b, int32_t b_is_matrix,
                             const void *out_grad, void *a_grad, void *b_grad, int32_t num_dims, int64_t cells,
                             int32_t dtype, int32_t flags, void *stream) {
    pict_status st = check_common(num_dims, cells, dtype);
    if (st) return st;
    PICT_CHECK_ARG(a && b && out_grad && a_grad && b_grad, "NULL tensor");
    if (cells == 0) return PICT_OK;
    const int kind = (a_is_matrix ? 1 : 0) + (b_is_matrix ? 2 : 0);
    hipStream_t s = (hipStream_t)stream;
    const unsigned g = grid_for((size_t)cells);
    PICT_MATMUL_DISPATCH(dtype, num_dims, kind,
                         hipLaunchKernelGGL((k_matmul_grad<R, D, K>), dim3(g), dim3(256), 0, s, (const R *)a,
                                            (const R *)b, (const R *)out_grad, (R *)a_grad, (R *)b_grad,
                                            (size_t)cells, (int)flags));
    return check_launch();
}

pict_status pict_vector_to_diag_matrix(const void *vectors, void *matrices, int32_t num_dims, int64_t cells,
                                       int32_t dtype, void *stream) {
    pict_status st = check_common(num_dims, cells, dtype);
    if (st) return st;
    PICT_CHECK_ARG(vectors && matrices, "NULL tensor");
    if (cells == 0) return PICT_OK;
    hipStream_t s = (hipStream_t)stream;
    const unsigned g = grid_for((size_t)cells);
    PICT_MATMUL_DISPATCH(dtype, num_dims, 0,
                         hipLaunchKernelGGL((k_vec_to_diag<R, D>), dim3(g), dim3(256), 0, s, (const R *)vectors,
                                            (R *)matrices, (size_t)cells));
    return check_launch();
}

pict_status pict_invert_matrix(const void *matrices, void *out, int32_t num_dims, int64_t cells, int32_t dtype,
                               void *stream) {
    pict_status st = check_common(num_dims, cells, dtype);
    if (st) return st;
    PICT_CHECK_ARG(matrices && out, "NULL tensor");
    if (cells == 0) return PICT_OK;
    hipStream_t s = (hipStream_t)stream;
    const unsigned g = grid_for((size_t)cells);
    PICT_MATMUL_DISPATCH(dtype, num_dims, 0,
                         hipLaunchKernelGGL((k_invert<R, D>), dim3(g), dim3(256), 0, s, (const R *)matrices, (R *)out,
                                            (size_t)cells));
    return check_launch();
}

}  // extern "C"
