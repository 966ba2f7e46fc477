// pict_sgs.hip -- spatial velocity gradients and the Smagorinsky sub-grid viscosity.
//
// Per cell (one thread, the stage kernels' tiles): for every velocity component, the gradient in
// physical coordinates (getBlockDataGradient, PISO_multiblock_cuda_kernel.cu:2982-3025), and from
// those the Smagorinsky viscosity C * Delta^2 * |S| (k_SGSviscosityIncompressibleSmagorinsky,
// :6898-6952).  Outputs are component-major global vectors; the Python layer cuts them into the
// per-block tensors the reference returns.
#include <hip/hip_runtime.h>

#include "pict_cells.hpp"
#include "pict_host.hpp"

namespace pict {

// Gradient of velocity component `comp` at a block cell (getBlockDataGradient :2982-3025): per axis
// the difference of the two neighbours' values over their distance in cells (2).  A face on a fixed
// boundary takes the boundary's value half a cell away (Dirichlet) or the cell's own value one cell
// away (Neumann); across a connection the neighbour block's cell gives the same component index (the
// reference does not remap components there, computeConnectedPos :327-347); a periodic face wraps.
// With a transform the row vector of computational derivatives is multiplied by Minv (matmul(v, M)
// of transformations.h:397-405: out_j = sum_i g_i Minv_ij).
template <typename R, int D>
__device__ __forceinline__ void velocity_gradient(const DomainK<R> &dom, const pict_block_desc *b, I4 pos, int comp,
                                                  R (&g)[D]) {
    pos.a[3] = comp;
    const R own = reinterpret_cast<const R *>(b->velocity)[flat4(pos, b->stride)];
    R cg[D];
#pragma unroll
    for (int dim = 0; dim < D; ++dim) {
        R dist = R(2), acc = R(0);
#pragma unroll
        for (int up = 0; up < 2; ++up) {
            const int face = 2 * dim + up;
            R v;
            if (at_bound(pos, face, b) && prescribed(face, b)) {
                const pict_boundary_desc *fb = &b->bounds[face];
                if (fb->velocity_bc == PICT_BC_DIRICHLET) {
                    I4 bp = pos;
                    bp.a[dim] = 0;
                    v = fixed_vel<R>(fb, bp, comp);
                    dist -= R(0.5);
                } else {
                    v = own;
                    dist -= R(1);
                }
            } else {
                const pict_block_desc *nb = b;
                I4 t = pos;
                if (at_bound(pos, face, b) && b->bounds[face].type == PICT_BOUNDARY_CONNECTED) {
                    nb = dom.blocks + b->bounds[face].connected_block;
                    t = connected_pos<D>(dom.blocks, pos, dim, &b->bounds[face], 0);
                } else if (at_bound(pos, face, b)) {
                    t.a[dim] = up ? 0 : b->size[dim] - 1;
                } else {
                    t.a[dim] += up ? 1 : -1;
                }
                v = reinterpret_cast<const R *>(nb->velocity)[flat4(t, nb->stride)];
            }
            acc += up ? v : -v;
        }
        cg[dim] = acc / dist;
    }
    if (b->has_transform) {
        const R *T = cell_T<R, D>(b, pos);
#pragma unroll
        for (int j = 0; j < D; ++j) {
            R s = 0;
#pragma unroll
            for (int i = 0; i < D; ++i) s += Tminv<R, D>(T, i, j) * cg[i];
            g[j] = s;
        }
    } else {
#pragma unroll
        for (int j = 0; j < D; ++j) g[j] = cg[j];
    }
}

// k_computeSpatialVelocityGradients :6445-6467: out[(comp * D + i) * N + g] = d u_comp / d x_i
template <typename R, int D>
__global__ __launch_bounds__(kCellsPerTile) void k_velocity_gradients(DomainK<R> dom, R *out) {
    const pict_block_desc *b;
    int bi, f;
    if (!tile_cell(dom, (int)blockIdx.x, b, bi, f)) return;
    const I4 pos = unflat(f, b);
    const size_t g0 = (size_t)b->global_offset + f;
#pragma unroll
    for (int comp = 0; comp < D; ++comp) {
        R g[D];
        velocity_gradient<R, D>(dom, b, pos, comp, g);
#pragma unroll
        for (int i = 0; i < D; ++i) out[(size_t)(comp * D + i) * dom.N + g0] = g[i];
    }
}

// k_SGSviscosityIncompressibleSmagorinsky :6898-6952: |S| = sqrt(2 S_ij S_ij) with S = (G + G^T)/2 (G_ij =
// d u_i / d x_j), the sum over i <= j with the off-diagonal terms doubled in the reference's order;
// Delta^2 = the largest squared column norm of the transform's M (1 without a transform);
// out[g] = C * Delta^2 * |S|
template <typename R, int D>
__global__ __launch_bounds__(kCellsPerTile) void k_sgs_smagorinsky(DomainK<R> dom, R coefficient, R *out) {
    const pict_block_desc *b;
    int bi, f;
    if (!tile_cell(dom, (int)blockIdx.x, b, bi, f)) return;
    const I4 pos = unflat(f, b);
    R G[D][D];
#pragma unroll
    for (int comp = 0; comp < D; ++comp) velocity_gradient<R, D>(dom, b, pos, comp, G[comp]);
    R s2 = 0;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = i; j < D; ++j) {
            R s = R(0.5) * (G[i][j] + G[j][i]);
            s *= s;
            if (i != j) s *= 2;
            s2 += s;
        }
    const R norm = sqrt(2 * s2);
    R delta = 1;
    if (b->has_transform) {
        const R *T = cell_T<R, D>(b, pos);  // M row-major in T[0 .. D*D)
        delta = 0;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            R col = 0;
#pragma unroll
            for (int r = 0; r < D; ++r) col += T[r * D + i] * T[r * D + i];
            delta = max(delta, col);
        }
    }
    out[(size_t)b->global_offset + f] = coefficient * delta * norm;
}

}  // namespace pict

using namespace pict;

extern "C" {

pict_status pict_compute_spatial_velocity_gradients(const pict_atlas *a, void *out, void *stream) {
    PICT_CHECK_ARG(a && out, "output missing");
    PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_velocity_gradients<R, D>), a->view<R>(), (R *)out));
    return check_launch();
}

pict_status pict_sgs_viscosity_smagorinsky(const pict_atlas *a, double coefficient, void *out, void *stream) {
    PICT_CHECK_ARG(a && out, "output missing");
    PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_sgs_smagorinsky<R, D>), a->view<R>(), (R)coefficient, (R *)out));
    return check_launch();
}

}  // extern "C"
