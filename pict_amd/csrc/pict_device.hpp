// pict_device.hpp -- device-side view of a domain and the per-cell helpers
// shared by the stage kernels (gfx950, wave64).
//
// Every helper restates one of the reference's device functions in
// tum-pbs/PICT extensions/PISO_multiblock_cuda_kernel.cu (cited per function).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pict_hip.h"

namespace pict {

constexpr int kCellsPerTile = 256;  // one workgroup = 256 consecutive cells of one block

struct I4 {
    int a[4];
};

// Kernel-argument view of a domain: scalars by value, tables in device memory.
template <typename R>
struct DomainK {
    int D, N, num_blocks, nnz;
    R viscosity;
    const pict_block_desc *blocks;  // device copy of the block table
    const int32_t *tile_block;      // per workgroup: block index
    const int32_t *tile_start;      // per workgroup: first flat cell in the block
    int num_tiles;
    int32_t *csr_row, *csr_index;
    R *C, *P, *A;
    R *vrhs, *vres, *prhs, *prhs_div, *pres;
    R *visc_grad, *C_grad, *P_grad, *A_grad;
    R *vrhs_grad, *vres_grad, *prhs_grad, *prhs_div_grad, *pres_grad;
    // passive scalar: channels, scalar viscosity ([1] or [C], null: the domain viscosity), global vectors [C*N]
    int channels, svisc_static;
    const R *svisc;
    R *svisc_grad, *srhs, *sres, *srhs_grad, *sres_grad;
};

__device__ __forceinline__ int flat4(const I4 &p, const int32_t *stride) {
    return p.a[0] + stride[1] * p.a[1] + stride[2] * p.a[2] + stride[3] * p.a[3];
}
__device__ __forceinline__ I4 unflat(int idx, const pict_block_desc *b) {
    I4 p;
    p.a[0] = idx % b->size[0];
    p.a[1] = (idx / b->stride[1]) % b->size[1];
    p.a[2] = (idx / b->stride[2]) % b->size[2];
    p.a[3] = 0;
    return p;
}
__device__ __forceinline__ int posmod(int a, int m) {
    int r = a % m;
    return r < 0 ? r + m : r;
}
// isAtBound :183-193
__device__ __forceinline__ bool at_bound(const I4 &p, int face, const pict_block_desc *b) {
    const int ax = face >> 1;
    return (face & 1) ? p.a[ax] == b->size[ax] - 1 : p.a[ax] == 0;
}
// isEmptyBound :159-166 (FIXED is the live prescribed type)
__device__ __forceinline__ bool prescribed(int face, const pict_block_desc *b) {
    return b->bounds[face].type == PICT_BOUNDARY_FIXED;
}
__device__ __forceinline__ int global_index(I4 p, const pict_block_desc *b) {
    p.a[3] = 0;
    return flat4(p, b->stride) + b->global_offset;
}

template <int D>
struct TS {
    static constexpr int n = 2 * D * D + 1;
};

template <typename R, int D>
__device__ __forceinline__ const R *cell_T(const pict_block_desc *b, I4 p) {
    p.a[3] = 0;
    return reinterpret_cast<const R *>(b->transform) + (size_t)flat4(p, b->stride) * TS<D>::n;
}
template <typename R, int D>
__device__ __forceinline__ const R *bound_T(const pict_boundary_desc *fb, I4 p) {
    p.a[3] = 0;
    return reinterpret_cast<const R *>(fb->transform) + (size_t)flat4(p, fb->stride) * TS<D>::n;
}
template <typename R, int D>
__device__ __forceinline__ R Tminv(const R *T, int i, int j) { return T[D * D + i * D + j]; }
template <typename R, int D>
__device__ __forceinline__ R Tdet(const R *T) { return T[2 * D * D]; }

// getFaceTransformPtr :874-889: a face's transform in the block's face transforms, laid out [D][(z+1)][(y+1)][x+1]
// (one more vertex along every spatial axis; an upper face sits one index further along its axis)
template <typename R, int D>
__device__ __forceinline__ const R *face_T(const pict_block_desc *b, I4 pos, int face) {
    const int dim = face >> 1;
    pos.a[dim] += face & 1;
    const size_t fx = b->size[0] + 1, fy = b->size[1] + (D > 1 ? 1 : 0), fz = b->size[2] + (D > 2 ? 1 : 0);
    const size_t idx = (size_t)pos.a[0] + fx * ((size_t)pos.a[1] + fy * ((size_t)pos.a[2] + fz * dim));
    return reinterpret_cast<const R *>(b->face_transform) + idx * TS<D>::n;
}

// getLaplaceCoefficientOrthogonalFace :1226-1241: det |Minv[axis]|^2 of the face's own transform (1 without face
// transforms)
template <typename R, int D>
__device__ __forceinline__ R laplace_orth_face(const pict_block_desc *b, const I4 &pos, int face) {
    if (!b->has_face_transform) return R(1);
    const R *T = face_T<R, D>(b, pos, face);
    const int ax = face >> 1;
    R s = 0;
#pragma unroll
    for (int j = 0; j < D; ++j) s += Tminv<R, D>(T, ax, j) * Tminv<R, D>(T, ax, j);
    return Tdet<R, D>(T) * s;
}

// computeConnectedPos :325-347
template <int D>
__device__ __forceinline__ I4 connected_pos(const pict_block_desc *blocks, I4 pos, int bdim,
                                            const pict_boundary_desc *cb, int border) {
    const pict_block_desc *ob = blocks + cb->connected_block;
    I4 o = pos;
    int ca = cb->axes[0] >> 1;
    o.a[ca] = (cb->axes[0] & 1) ? ob->size[ca] - 1 - border : border;
    if (D > 1) {
        int ax = (bdim + 1) % D;
        ca = cb->axes[1] >> 1;
        o.a[ca] = (cb->axes[1] & 1) ? ob->size[ca] - 1 - pos.a[ax] : pos.a[ax];
        if (D > 2) {
            ax = (bdim + 2) % D;
            ca = cb->axes[2] >> 1;
            o.a[ca] = (cb->axes[2] & 1) ? ob->size[ca] - 1 - pos.a[ax] : pos.a[ax];
        }
    }
    return o;
}
// computeConnectedPosWithChannel :351-360
template <int D>
__device__ __forceinline__ I4 connected_pos_ch(const pict_block_desc *blocks, I4 pos, int bdim,
                                               const pict_boundary_desc *cb) {
    I4 o = connected_pos<D>(blocks, pos, bdim, cb, 0);
    o.a[3] = cb->axes[posmod(pos.a[3] - bdim, D)] >> 1;
    return o;
}

// Neighbour across a non-prescribed face with the channel remapped across
// connections (PISO_build_matrix :3716-3729, PISO_build_pressure_matrix :4835-4845).
template <int D>
__device__ __forceinline__ I4 neighbor_cell(const pict_block_desc *blocks, const pict_block_desc *b,
                                            I4 pos, int face, int self, int &nb) {
    const int dim = face >> 1, up = face & 1;
    I4 t = pos;
    t.a[3] = dim;
    nb = self;
    const bool ab = at_bound(pos, face, b);
    if (ab && b->bounds[face].type == PICT_BOUNDARY_CONNECTED) {
        nb = b->bounds[face].connected_block;
        t = connected_pos_ch<D>(blocks, t, dim, &b->bounds[face]);
    } else if (ab && b->bounds[face].type == PICT_BOUNDARY_PERIODIC) {
        t.a[dim] = up ? 0 : b->size[dim] - 1;
    } else {
        t.a[dim] = pos.a[dim] + (up ? 1 : -1);
    }
    return t;
}

// velocity component d of a block cell from the block field or a global vector (:577-593)
template <typename R>
__device__ __forceinline__ R vel_comp(const DomainK<R> &dom, const pict_block_desc *b, I4 p, int d,
                                      const R *vg) {
    p.a[3] = d;
    if (!vg) return reinterpret_cast<const R *>(b->velocity)[flat4(p, b->stride)];
    return vg[(size_t)d * dom.N + global_index(p, b)];
}

// getContravariantComponent :665-678 (p.w = component)
template <typename R, int D>
__device__ __forceinline__ R contravariant(const DomainK<R> &dom, const pict_block_desc *b, I4 p,
                                           const R *vg) {
    const int comp = p.a[3];
    if (b->has_transform) {
        const R *T = cell_T<R, D>(b, p);
        R s = 0;
#pragma unroll
        for (int j = 0; j < D; ++j) s += Tminv<R, D>(T, comp, j) * vel_comp(dom, b, p, j, vg);
        return Tdet<R, D>(T) * s;
    }
    return vel_comp(dom, b, p, comp, vg);
}

// FixedBoundary velocity, getFixedBoundaryData :2419-2437
template <typename R>
__device__ __forceinline__ R fixed_vel(const pict_boundary_desc *fb, I4 p, int d) {
    if (fb->velocity_static) return reinterpret_cast<const R *>(fb->velocity)[d];
    p.a[3] = d;
    return reinterpret_cast<const R *>(fb->velocity)[flat4(p, fb->stride)];
}

// getContravariantComponentBoundaryFixed :690-710 (p on the boundary grid, w = component)
template <typename R, int D>
__device__ __forceinline__ R fixed_contravariant(const pict_boundary_desc *fb, I4 p) {
    const int comp = p.a[3];
    if (fb->has_transform) {
        const R *T = bound_T<R, D>(fb, p);
        R s = 0;
        if (fb->velocity_bc == PICT_BC_DIRICHLET) {
#pragma unroll
            for (int j = 0; j < D; ++j) s += Tminv<R, D>(T, comp, j) * fixed_vel<R>(fb, p, j);
        }
        return Tdet<R, D>(T) * s;
    }
    return fixed_vel<R>(fb, p, comp);
}

// getLaplaceCoefficientOrthogonal :1209-1224
template <typename R, int D>
__device__ __forceinline__ R laplace_orth(const pict_block_desc *b, I4 p) {
    if (!b->has_transform) return R(1);
    const R *T = cell_T<R, D>(b, p);
    R s = 0;
#pragma unroll
    for (int j = 0; j < D; ++j) s += Tminv<R, D>(T, p.a[3], j) * Tminv<R, D>(T, p.a[3], j);
    return Tdet<R, D>(T) * s;
}
// getLaplaceCoefficientOrthogonalBoundary :1454-1466
template <typename R, int D>
__device__ __forceinline__ R laplace_orth_fixed(const pict_boundary_desc *fb, I4 p) {
    if (!fb->has_transform) return R(1);
    const R *T = bound_T<R, D>(fb, p);
    R s = 0;
#pragma unroll
    for (int j = 0; j < D; ++j) s += Tminv<R, D>(T, p.a[3], j) * Tminv<R, D>(T, p.a[3], j);
    return Tdet<R, D>(T) * s;
}

// getViscosityBlock :1802-1821
template <typename R>
__device__ __forceinline__ R viscosity(const DomainK<R> &dom, const pict_block_desc *b, I4 p) {
    if (!b->viscosity) return dom.viscosity;
    if (b->viscosity_static) return reinterpret_cast<const R *>(b->viscosity)[0];
    p.a[3] = 0;
    return reinterpret_cast<const R *>(b->viscosity)[flat4(p, b->stride)];
}

// getViscosity(forPassiveScalar) :1788-1800 (the scalar viscosity is global, getViscosityBlock :1803-1807)
template <typename R>
__device__ __forceinline__ R scalar_viscosity(const DomainK<R> &dom, int ch) {
    if (!dom.svisc) return dom.viscosity;
    return dom.svisc[dom.svisc_static ? 0 : ch];
}
// FixedBoundary passive scalar of channel ch, getFixedBoundaryData :2419-2437 (p on the boundary grid)
template <typename R>
__device__ __forceinline__ R fixed_scalar(const pict_boundary_desc *fb, I4 p, int ch) {
    if (fb->scalar_static) return reinterpret_cast<const R *>(fb->scalar)[ch];
    p.a[3] = ch;
    return reinterpret_cast<const R *>(fb->scalar)[flat4(p, fb->stride)];
}
// getFixedBoundaryType (PASSIVE_SCALAR) :2360-2370
__device__ __forceinline__ int fixed_scalar_bc(const pict_boundary_desc *fb, int ch) {
    return ((fb->scalar_bc_mask >> (ch & 31)) & 1) ? PICT_BC_NEUMANN : PICT_BC_DIRICHLET;
}

// getBlockVelocitySource :2241-2252
template <typename R>
__device__ __forceinline__ R velocity_source(const pict_block_desc *b, const I4 &p) {
    if (!b->velocity_source) return R(0);
    if (b->velocity_source_static) return reinterpret_cast<const R *>(b->velocity_source)[p.a[3]];
    return reinterpret_cast<const R *>(b->velocity_source)[flat4(p, b->stride)];
}

template <typename R, int D>
__device__ __forceinline__ R cell_det(const pict_block_desc *b, const I4 &p) {
    return b->has_transform ? Tdet<R, D>(cell_T<R, D>(b, p)) : R(1);
}

// computeFluxesNDLoop :1552-1630 -- face fluxes, NOT multiplied by the face sign
template <typename R, int D>
__device__ __forceinline__ void fluxes_nd(const DomainK<R> &dom, const pict_block_desc *b, const I4 &pos,
                                          R *fl, const R *vg) {
#pragma unroll
    for (int face = 0; face < 2 * D; ++face) {
        const int dim = face >> 1, up = face & 1;
        I4 t = pos;
        t.a[3] = dim;
        const R velC = contravariant<R, D>(dom, b, t, vg);
        if (at_bound(pos, face, b)) {
            const pict_boundary_desc *bd = &b->bounds[face];
            if (bd->type == PICT_BOUNDARY_FIXED) {
                t.a[dim] = 0;
                fl[face] = fixed_contravariant<R, D>(bd, t);
            } else if (bd->type == PICT_BOUNDARY_CONNECTED) {
                const pict_block_desc *ob = dom.blocks + bd->connected_block;
                I4 o = connected_pos<D>(dom.blocks, t, dim, bd, 0);
                o.a[3] = bd->axes[0] >> 1;
                R velN = contravariant<R, D>(dom, ob, o, vg);
                if ((bd->axes[0] & 1) == up) velN = -velN;
                fl[face] = (velN + velC) * R(0.5);
            } else if (bd->type == PICT_BOUNDARY_PERIODIC) {
                t.a[dim] = up ? 0 : b->size[dim] - 1;
                fl[face] = (contravariant<R, D>(dom, b, t, vg) + velC) * R(0.5);
            } else {
                fl[face] = 0;
            }
        } else {
            t.a[dim] = pos.a[dim] + (up ? 1 : -1);
            fl[face] = (contravariant<R, D>(dom, b, t, vg) + velC) * R(0.5);
        }
    }
}

// Boundary source terms of kPISO_build_advection_RHS :4300-4330 and
// PISO_build_pressure_rhs :5166-5188 for FIXED faces at the block border.
template <typename R, int D>
__device__ __forceinline__ R fixed_boundary_source(const DomainK<R> &dom, const pict_block_desc *b,
                                                   const I4 &pos, int d) {
    R s = 0;
#pragma unroll
    for (int face = 0; face < 2 * D; ++face) {
        if (!at_bound(pos, face, b)) continue;
        const pict_boundary_desc *fb = &b->bounds[face];
        if (fb->type != PICT_BOUNDARY_FIXED) continue;
        const int ax = face >> 1;
        const R normal = (face & 1) ? R(1) : R(-1);
        I4 t = pos;
        t.a[3] = d;
        t.a[ax] = 0;
        const R vel = fixed_vel<R>(fb, t, d);
        t.a[3] = ax;
        const R alpha = laplace_orth_fixed<R, D>(fb, t);
        const R flux = fixed_contravariant<R, D>(fb, t) * normal;
        s -= vel * flux;
        const R visc = viscosity(dom, b, pos);
        const int slip = fb->velocity_bc == PICT_BC_DIRICHLET ? 0 : 1;
        s += vel * R(1 - slip) * visc * R(2) * alpha;
    }
    return s;
}

// getPressureAtWithBounds :2160-2239 (zeroBound = false)
template <typename R, int D>
__device__ __forceinline__ R pressure_at(const DomainK<R> &dom, const pict_block_desc *b, const I4 &pos) {
    I4 t = pos;
    t.a[3] = 0;
#pragma unroll
    for (int dim = 0; dim < D; ++dim) {
        if (pos.a[dim] < 0 || pos.a[dim] >= b->size[dim]) {
            const bool lower = pos.a[dim] < 0;
            const pict_boundary_desc *bd = &b->bounds[2 * dim + (lower ? 0 : 1)];
            if (bd->type == PICT_BOUNDARY_FIXED) {
                t.a[dim] = lower ? 0 : b->size[dim] - 1;
                return reinterpret_cast<const R *>(b->pressure)[flat4(t, b->stride)];
            }
            if (bd->type == PICT_BOUNDARY_CONNECTED) {
                const pict_block_desc *ob = dom.blocks + bd->connected_block;
                const int border = lower ? -(pos.a[dim] + 1) : pos.a[dim] - b->size[dim];
                I4 o = connected_pos<D>(dom.blocks, t, dim, bd, border);
                return reinterpret_cast<const R *>(ob->pressure)[flat4(o, ob->stride)];
            }
            if (bd->type == PICT_BOUNDARY_PERIODIC) {
                t.a[dim] += lower ? b->size[dim] : -b->size[dim];
                return reinterpret_cast<const R *>(b->pressure)[flat4(t, b->stride)];
            }
            return R(0);
        }
    }
    return reinterpret_cast<const R *>(b->pressure)[flat4(t, b->stride)];
}

// getPressureGradient :801-834, gradient in physical space (g^T Minv)
template <typename R, int D>
__device__ __forceinline__ void pressure_gradient(const DomainK<R> &dom, const pict_block_desc *b,
                                                  const I4 &pos, R *g) {
    R gr[D];
#pragma unroll
    for (int dim = 0; dim < D; ++dim) {
        I4 t = pos;
        t.a[3] = 0;
        R fac = R(0.5);
        if (!(pos.a[dim] == 0 && prescribed(2 * dim, b))) t.a[dim] = pos.a[dim] - 1;
        else fac = R(1);
        const R vN = pressure_at<R, D>(dom, b, t);
        if (!(pos.a[dim] == b->size[dim] - 1 && prescribed(2 * dim + 1, b))) t.a[dim] = pos.a[dim] + 1;
        else {
            t.a[dim] = pos.a[dim];
            fac = R(1);
        }
        const R vP = pressure_at<R, D>(dom, b, t);
        gr[dim] = (vP - vN) * fac;
    }
    if (b->has_transform) {
        const R *T = cell_T<R, D>(b, pos);
#pragma unroll
        for (int j = 0; j < D; ++j) {
            R s = 0;
#pragma unroll
            for (int i = 0; i < D; ++i) s += Tminv<R, D>(T, i, j) * gr[i];
            g[j] = s;
        }
    } else {
#pragma unroll
        for (int j = 0; j < D; ++j) g[j] = gr[j];
    }
}

// Unsorted row entries of a cell: slot 0 = diagonal, slot face+1 = neighbour
// across `face` (-1 for prescribed faces); PISO_build_matrix :3663-3772.
template <typename R, int D>
__device__ __forceinline__ void row_slots(const DomainK<R> &dom, const pict_block_desc *b, int bi, const I4 &pos,
                                          int *idx) {
#pragma unroll
    for (int k = 0; k < 7; ++k) idx[k] = -1;
    idx[0] = global_index(pos, b);
#pragma unroll
    for (int face = 0; face < 2 * D; ++face) {
        if (at_bound(pos, face, b) && prescribed(face, b)) continue;
        int nb;
        const I4 t = neighbor_cell<D>(dom.blocks, b, pos, face, bi, nb);
        idx[face + 1] = global_index(t, dom.blocks + nb);
    }
}

// findLowestColumnIndex :3479-3489
__device__ __forceinline__ int lowest_slot(const int *idx) {
    int v = 0x7fffffff, k = -1;
#pragma unroll
    for (int i = 0; i < 7; ++i)
        if (idx[i] >= 0 && idx[i] < v) {
            v = idx[i];
            k = i;
        }
    return k;
}

// Resolve the workgroup's (block, first cell) from the launch tables.
template <typename R>
__device__ __forceinline__ bool tile_cell(const DomainK<R> &dom, int tile, const pict_block_desc *&b, int &bi,
                                          int &f) {
    bi = dom.tile_block[tile];
    b = dom.blocks + bi;
    f = dom.tile_start[tile] + (int)threadIdx.x;
    return f < b->stride[3];
}
template <typename R>
__device__ __forceinline__ bool tile_cell(const DomainK<R> &dom, const pict_block_desc *&b, int &bi, int &f) {
    return tile_cell(dom, (int)blockIdx.x, b, bi, f);
}

}  // namespace pict
