// pict_metrics.hip -- cell and face transformation metrics from vertex coordinates
// (CoordsToTransforms grid_gen.cu:284-381, CoordsToFaceTransforms grid_gen.cu:383-514).
// Setup-only kernels: one thread per cell (face position).
#include <hip/hip_runtime.h>

#include "pict_host.hpp"

namespace pict {

template <typename R, int D>
__device__ __forceinline__ R det_n(const R (&M)[D][D]) {
    if constexpr (D == 1) return M[0][0];
    else if constexpr (D == 2) return M[0][0] * M[1][1] - M[0][1] * M[1][0];
    else
        return M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
               M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
}
// adjugate / determinant (transformations.h:476-560)
template <typename R, int D>
__device__ __forceinline__ void inv_n(const R (&M)[D][D], R det, R *out) {
    if constexpr (D == 1) {
        out[0] = R(1) / M[0][0];
    } else if constexpr (D == 2) {
        out[0] = M[1][1] / det; out[1] = -M[0][1] / det;
        out[2] = -M[1][0] / det; out[3] = M[0][0] / det;
    } else {
        out[0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) / det;
        out[1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) / det;
        out[2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) / det;
        out[3] = (M[1][2] * M[2][0] - M[1][0] * M[2][2]) / det;
        out[4] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) / det;
        out[5] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) / det;
        out[6] = (M[1][0] * M[2][1] - M[1][1] * M[2][0]) / det;
        out[7] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) / det;
        out[8] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) / det;
    }
}

template <typename R, int D>
__device__ __forceinline__ void write_T(const R (&M)[D][D], R *T) {
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int k = 0; k < D; ++k) T[i * D + k] = M[i][k];
    const R det = det_n<R, D>(M);
    inv_n<R, D>(M, det, T + D * D);
    T[2 * D * D] = det;
}

// k_CoordsToTransforms grid_gen.cu:284-339: M[i][k] = centre(+k face) - centre(-k face), component i
template <typename R, int D>
__global__ void k_coords_to_T(const R *coords, int VX, int VY, int VZ, R *out) {
    const int CX = VX - 1, CY = D > 1 ? VY - 1 : 1, CZ = D > 2 ? VZ - 1 : 1;
    const int cells = CX * CY * CZ;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= cells) return;
    const int x = idx % CX, y = (idx / CX) % CY, z = idx / (CX * CY);
    const size_t cs = (size_t)VX * VY * VZ;
    R fc[2 * D][D];
#pragma unroll
    for (int f = 0; f < 2 * D; ++f)
#pragma unroll
        for (int k = 0; k < D; ++k) fc[f][k] = 0;
#pragma unroll
    for (int v = 0; v < (1 << D); ++v) {
        const int vx = x + (v & 1), vy = y + ((v & 2) != 0), vz = z + ((v & 4) != 0);
        const size_t vi = vx + (size_t)VX * vy + (size_t)VX * VY * vz;
#pragma unroll
        for (int f = 0; f < D; ++f) {
            const int up = (v & (1 << f)) != 0;
#pragma unroll
            for (int k = 0; k < D; ++k) fc[f * 2 + up][k] += coords[k * cs + vi];
        }
    }
    const R norm = R(1.0) / R(1 << (D - 1));
    R M[D][D];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int k = 0; k < D; ++k) M[i][k] = fc[k * 2 + 1][i] * norm - fc[k * 2][i] * norm;
    write_T<R, D>(M, out + (size_t)idx * (2 * D * D + 1));
}

// k_CoordsToFaceTransforms grid_gen.cu:383-479 (staggered, one thread per vertex position)
template <typename R, int D>
__global__ void k_coords_to_face_T(const R *coords, int SX, int SY, int SZ, R *out) {
    const int size[3] = {SX, D > 1 ? SY : 1, D > 2 ? SZ : 1};
    const int total = size[0] * size[1] * size[2];
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int pos[3] = {idx % size[0], (idx / size[0]) % size[1], idx / (size[0] * size[1])};
    const size_t cs = (size_t)total;
    constexpr int nfv = 1 << (D - 1);
    constexpr int nev = D >= 2 ? 1 << (D - 2) : 1;
    auto vidx = [&](const int *p) { return (size_t)p[0] + (size_t)size[0] * p[1] + (size_t)size[0] * size[1] * p[2]; };
#pragma unroll
    for (int dim = 0; dim < D; ++dim) {
        bool skip = false;
#pragma unroll
        for (int i = 1; i < D; ++i) {
            const int td = (dim + i) % D;
            if (pos[td] == size[td] - 1) skip = true;
        }
        if (skip) continue;
        R M[D][D];
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int k = 0; k < D; ++k) M[i][k] = 0;
        R nw = R(0.5);
#pragma unroll
        for (int up = 0; up < 2; ++up) {
            const int fs = up * 2 - 1;
            const bool atb = up ? pos[dim] == size[dim] - 1 : pos[dim] == 0;
            int fp[3] = {pos[0], pos[1], pos[2]};
            if (!atb) fp[dim] += fs;
            else nw = R(1.0);
#pragma unroll
            for (int v = 0; v < nfv; ++v) {
                int vp[3] = {fp[0], fp[1], fp[2]};
#pragma unroll
                for (int i = 1; i < D; ++i) vp[(dim + i) % D] += (v & i) != 0;
#pragma unroll
                for (int k = 0; k < D; ++k) M[k][dim] += coords[k * cs + vidx(vp)] * R(fs);
            }
        }
        nw /= R(nfv);
#pragma unroll
        for (int k = 0; k < D; ++k) M[k][dim] *= nw;
        const R tn = R(1.0) / R(nev);
#pragma unroll
        for (int kk = 1; kk < D; ++kk) {
            const int td = (dim + kk) % D;
#pragma unroll
            for (int v = 0; v < nfv; ++v) {
                int vp[3] = {pos[0], pos[1], pos[2]};
#pragma unroll
                for (int i = 1; i < D; ++i) vp[(dim + i) % D] += (v & i) != 0;
                const int tfs = (v & kk) ? 1 : -1;
#pragma unroll
                for (int k = 0; k < D; ++k) M[k][td] += coords[k * cs + vidx(vp)] * R(tfs);
            }
#pragma unroll
            for (int k = 0; k < D; ++k) M[k][td] *= tn;
        }
        write_T<R, D>(M, out + ((size_t)dim * cs + idx) * (2 * D * D + 1));
    }
}

}  // namespace pict

using namespace pict;

extern "C" {

pict_status pict_coords_to_transforms(const void *coords, int32_t D, const int32_t vs[3], int32_t dtype, void *out,
                                      void *stream) {
    PICT_CHECK_ARG(coords && out && vs, "NULL argument");
    PICT_CHECK_ARG(D >= 1 && D <= 3, "invalid dims");
    const int VX = vs[0], VY = D > 1 ? vs[1] : 2, VZ = D > 2 ? vs[2] : 2;
    PICT_CHECK_ARG(VX >= 2 && VY >= 2 && VZ >= 2, "vertex grid too small");
    const int cells = (VX - 1) * (D > 1 ? VY - 1 : 1) * (D > 2 ? VZ - 1 : 1);
    const dim3 grid((cells + 255) / 256), block(256);
    hipStream_t s = (hipStream_t)stream;
#define PICT_T_LAUNCH(R_, D_) \
    hipLaunchKernelGGL((k_coords_to_T<R_, D_>), grid, block, 0, s, (const R_ *)coords, VX, D_ > 1 ? VY : 1, D_ > 2 ? VZ : 1, (R_ *)out)
    if (dtype == PICT_DTYPE_F32) {
        if (D == 1) PICT_T_LAUNCH(float, 1); else if (D == 2) PICT_T_LAUNCH(float, 2); else PICT_T_LAUNCH(float, 3);
    } else {
        if (D == 1) PICT_T_LAUNCH(double, 1); else if (D == 2) PICT_T_LAUNCH(double, 2); else PICT_T_LAUNCH(double, 3);
    }
#undef PICT_T_LAUNCH
    PICT_HIP_TRY(hipGetLastError());
    return PICT_OK;
}

pict_status pict_coords_to_face_transforms(const void *coords, int32_t D, const int32_t vs[3], int32_t dtype,
                                           void *out, void *stream) {
    PICT_CHECK_ARG(coords && out && vs, "NULL argument");
    PICT_CHECK_ARG(D >= 2 && D <= 3, "face transforms need 2 or 3 dims");
    const int SX = vs[0], SY = vs[1], SZ = D > 2 ? vs[2] : 1;
    const int total = SX * SY * SZ;
    const dim3 grid((total + 255) / 256), block(256);
    hipStream_t s = (hipStream_t)stream;
    if (dtype == PICT_DTYPE_F32) {
        if (D == 2) hipLaunchKernelGGL((k_coords_to_face_T<float, 2>), grid, block, 0, s, (const float *)coords, SX, SY, SZ, (float *)out);
        else hipLaunchKernelGGL((k_coords_to_face_T<float, 3>), grid, block, 0, s, (const float *)coords, SX, SY, SZ, (float *)out);
    } else {
        if (D == 2) hipLaunchKernelGGL((k_coords_to_face_T<double, 2>), grid, block, 0, s, (const double *)coords, SX, SY, SZ, (double *)out);
        else hipLaunchKernelGGL((k_coords_to_face_T<double, 3>), grid, block, 0, s, (const double *)coords, SX, SY, SZ, (double *)out);
    }
    PICT_HIP_TRY(hipGetLastError());
    return PICT_OK;
}

}  // extern "C"
