// pict_nonortho.hpp -- non-orthogonal correction terms of the Laplacian on transformed grids (flags of
// PISO_multiblock_cuda.h:9-13), per cell, for the stage kernels.
//
// The cross-derivative terms alpha_cd d/dxi_d (d/dxi_c p), alpha = det Minv Minv^T, are split between the matrix
// and an explicit right-hand side: the face's tangential derivative is a difference of two corner values, each the
// mean of the cells around the corner (found by walking around it through connections and periodic wraps); the
// corner cells that are the centre cell or its face neighbours go into the matrix (CENTER_MATRIX, DIRECT_MATRIX),
// the others are read from the previous iterate (DIRECT_RHS, DIAGONAL_RHS).  Each function restates one of the
// reference's device functions in tum-pbs/PICT extensions/PISO_multiblock_cuda_kernel.cu (cited per function),
// including its quirks: corner walks and neighbour reads across a block connection land on the connected block's
// second cell layer (border offset 1), and the end cells of an axis with a prescribed face get zero interpolated
// cross coefficients on both faces of that axis.
#pragma once

#include <hip/hip_runtime.h>

#include "pict_device.hpp"

namespace pict {

// Field a corner or neighbour value is read from (the GridDataType values the reference's stages pass,
// domain_structs_gpu.h:28-65).
enum GridData {
    GD_VELOCITY = 0, GD_VELOCITY_RESULT = 1, GD_PRESSURE_RESULT = 2, GD_IS_FIXED = 3,
    GD_SCALAR = 4, GD_SCALAR_RESULT = 5  // passive scalar: p.w is the channel
};
__host__ __device__ constexpr bool gd_scalar(int type) { return type == GD_SCALAR || type == GD_SCALAR_RESULT; }

constexpr int kNonOrthoMatrix = PICT_NON_ORTHO_DIRECT_MATRIX | PICT_NON_ORTHO_CENTER_MATRIX;
constexpr int kNonOrthoRhs = PICT_NON_ORTHO_DIRECT_RHS | PICT_NON_ORTHO_DIAGONAL_RHS;

__device__ __forceinline__ int face_sign(int face) { return (face & 1) ? 1 : -1; }
__device__ __forceinline__ bool i4_eq(const I4 &a, const I4 &b) {
    return a.a[0] == b.a[0] && a.a[1] == b.a[1] && a.a[2] == b.a[2] && a.a[3] == b.a[3];
}
// getLaplaceCoefficientCenterIndex :1253-1255
__host__ __device__ constexpr int lap_ci(int c1, int c2) { return c1 + c2 + ((c1 >> 1) | (c2 >> 1)); }

// dot of Minv rows r1 and r2, ascending sum (transformations.h dot)
template <typename R, int D>
__device__ __forceinline__ R minv_dot(const R *T, int r1, int r2) {
    R s = 0;
#pragma unroll
    for (int j = 0; j < D; ++j) s += Tminv<R, D>(T, r1, j) * Tminv<R, D>(T, r2, j);
    return s;
}
// getLaplaceCoefficientsFull :1261-1287
template <typename R, int D>
__device__ __forceinline__ void lap_full(const pict_block_desc *b, const I4 &p, R *out) {
    const R *T = b->has_transform ? cell_T<R, D>(b, p) : nullptr;
#pragma unroll
    for (int c1 = 0; c1 < D; ++c1)
#pragma unroll
        for (int c2 = 0; c2 <= c1; ++c2)
            out[lap_ci(c1, c2)] = T ? Tdet<R, D>(T) * minv_dot<R, D>(T, c1, c2) : (c1 == c2 ? R(1) : R(0));
}
// getLaplaceCoefficientsSingleAxis :1289-1313
template <typename R, int D>
__device__ __forceinline__ void lap_single_axis(const pict_block_desc *b, const I4 &p, int axis, R *out) {
    const R *T = b->has_transform ? cell_T<R, D>(b, p) : nullptr;
#pragma unroll
    for (int c1 = 0; c1 < D; ++c1) out[c1] = T ? Tdet<R, D>(T) * minv_dot<R, D>(T, c1, axis) : (c1 == axis ? R(1) : R(0));
}
// getLaplaceCoefficientsNeighbourBoundary :1468-1490 (FixedBoundary)
template <typename R, int D>
__device__ __forceinline__ void lap_neighbour_boundary(const pict_boundary_desc *fb, I4 p, int axis, R *out) {
    p.a[3] = 0;
    p.a[axis] = 0;
    const R *T = fb->has_transform ? bound_T<R, D>(fb, p) : nullptr;
#pragma unroll
    for (int c1 = 0; c1 < D; ++c1) out[c1] = T ? Tdet<R, D>(T) * minv_dot<R, D>(T, c1, axis) : (c1 == axis ? R(1) : R(0));
}
// computeConnectedDir :314-323
template <int D>
__device__ __forceinline__ int connected_dir(int dir, int bdim, const pict_boundary_desc *cb) {
    return cb->axes[posmod((dir >> 1) - bdim, D)] ^ (dir & 1);
}

// A neighbour cell with the axis mapping of the connection crossed (resolveNeighborCell :431-476, block cells).
struct NCell {
    int bi;
    I4 pos;
    int map[3];
};
template <int D>
__device__ __forceinline__ NCell resolve_neighbor(const pict_block_desc *blocks, int bi, const I4 &pos, int dir) {
    const pict_block_desc *b = blocks + bi;
    NCell n;
    n.bi = bi;
    n.pos = pos;
    n.map[0] = 0;
    n.map[1] = 2;
    n.map[2] = 4;
    const int axis = dir >> 1;
    if (at_bound(pos, dir, b)) {
        const pict_boundary_desc *bd = &b->bounds[dir];
        if (bd->type == PICT_BOUNDARY_CONNECTED) {
            n.bi = bd->connected_block;
            n.pos = connected_pos<D>(blocks, pos, axis, bd, 0);
#pragma unroll
            for (int d = 0; d < D; ++d) n.map[d] = connected_dir<D>(n.map[d], axis, bd);
        } else if (bd->type == PICT_BOUNDARY_PERIODIC) {
            n.pos.a[axis] = (dir & 1) ? 0 : b->size[axis] - 1;
        }
    } else {
        n.pos.a[axis] += face_sign(dir);
    }
    return n;
}
// getLaplaceCoefficientsSingleAxisNeighbor(NeighborCellInfo) :1415-1452
template <typename R, int D>
__device__ __forceinline__ void lap_single_axis_neighbor(const pict_block_desc *blocks, const NCell &n, int dir,
                                                         R *out) {
    const int axis = dir >> 1;
    const pict_block_desc *nb = blocks + n.bi;
#pragma unroll
    for (int c1 = 0; c1 < D; ++c1) out[c1] = 0;
    if (nb->has_transform) {
        I4 q = n.pos;
        q.a[3] = 0;
        const R *T = cell_T<R, D>(nb, q);
        const int am = n.map[axis] >> 1;
#pragma unroll
        for (int c1 = 0; c1 < D; ++c1) out[c1] = Tdet<R, D>(T) * minv_dot<R, D>(T, n.map[c1] >> 1, am);
    } else {
        out[axis] = 1;
    }
}

// getBlockData :2320-2331 (p.w = component)
template <typename R>
__device__ __forceinline__ R block_data(const DomainK<R> &dom, int bi, const I4 &p, int type) {
    const pict_block_desc *b = dom.blocks + bi;
    switch (type) {
        case GD_VELOCITY: return reinterpret_cast<const R *>(b->velocity)[flat4(p, b->stride)];
        case GD_VELOCITY_RESULT: return dom.vres[(size_t)p.a[3] * dom.N + global_index(p, b)];
        case GD_PRESSURE_RESULT: return dom.pres[(size_t)p.a[3] * dom.N + global_index(p, b)];
        case GD_SCALAR: return reinterpret_cast<const R *>(b->scalar)[flat4(p, b->stride)];
        case GD_SCALAR_RESULT: return dom.sres[(size_t)p.a[3] * dom.N + global_index(p, b)];
        default: return R(0);
    }
}
// getFixedBoundaryData :2376-2446 for FIXED faces (the pressure of a FixedBoundary is a static zero)
template <typename R>
__device__ __forceinline__ R fixed_data(const pict_block_desc *b, I4 p, int face, int type) {
    if (type == GD_IS_FIXED) return R(1);
    const pict_boundary_desc *fb = &b->bounds[face];
    if (fb->type != PICT_BOUNDARY_FIXED || type == GD_PRESSURE_RESULT) return R(0);
    if (gd_scalar(type)) {
        p.a[face >> 1] = 0;
        return fixed_scalar<R>(fb, p, p.a[3]);
    }
    if (fb->velocity_static) return reinterpret_cast<const R *>(fb->velocity)[p.a[3]];
    p.a[face >> 1] = 0;
    return reinterpret_cast<const R *>(fb->velocity)[flat4(p, fb->stride)];
}
// getFixedBoundaryType :2351-2374 (ch: the passive-scalar channel, pos.w in the reference)
__device__ __forceinline__ int fixed_type(const pict_block_desc *b, int face, int type, int ch = 0) {
    const pict_boundary_desc *fb = &b->bounds[face];
    if (fb->type != PICT_BOUNDARY_FIXED) return PICT_BC_DIRICHLET;
    if (type == GD_VELOCITY || type == GD_VELOCITY_RESULT) return fb->velocity_bc;
    if (type == GD_PRESSURE_RESULT) return PICT_BC_NEUMANN;
    if (gd_scalar(type)) return fixed_scalar_bc(fb, ch);
    return PICT_BC_DIRICHLET;
}

// One step from (bi, t) across `dir`; returns false at a FIXED face (t, bi unchanged) -- the moves of
// getBlockDataNeighbor :2518-2557 and getBlockDataNeighborDiagonal :2614-2656.
template <int D>
__device__ __forceinline__ bool step_data(const pict_block_desc *blocks, int &bi, I4 &t, int dir) {
    const pict_block_desc *b = blocks + bi;
    const int dim = dir >> 1;
    if (at_bound(t, dir, b)) {
        const pict_boundary_desc *bd = &b->bounds[dir];
        if (bd->type == PICT_BOUNDARY_CONNECTED) {
            t = connected_pos<D>(blocks, t, dim, bd, 1);
            bi = bd->connected_block;
        } else if (bd->type == PICT_BOUNDARY_PERIODIC) {
            t.a[dim] = (dir & 1) ? 0 : b->size[dim] - 1;
        } else {
            return false;
        }
    } else {
        t.a[dim] += face_sign(dir);
    }
    return true;
}
// getBlockDataNeighbor :2518-2557 (zeroBound false)
template <typename R, int D>
__device__ __forceinline__ R data_neighbor(const DomainK<R> &dom, int bi, const I4 &pos, int dir, int type) {
    int tb = bi;
    I4 t = pos;
    if (!step_data<D>(dom.blocks, tb, t, dir)) return fixed_data<R>(dom.blocks + bi, pos, dir, type);
    return block_data(dom, tb, t, type);
}
// getBlockDataNeighborDiagonal :2614-2656: two steps, a prescribed first direction taken last
template <typename R, int D>
__device__ __forceinline__ R data_neighbor_diag(const DomainK<R> &dom, int bi, const I4 &pos, int dir1, int dir2,
                                                int type) {
    const bool d1e = prescribed(dir1, dom.blocks + bi);
    const int dirs[2] = {d1e ? dir2 : dir1, d1e ? dir1 : dir2};
    int tb = bi;
    I4 t = pos;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int cb = tb;
        const I4 ct = t;
        if (!step_data<D>(dom.blocks, tb, t, dirs[i])) return fixed_data<R>(dom.blocks + cb, ct, dirs[i], type);
    }
    return block_data(dom, tb, t, type);
}

// A walker of the corner walk (CycleDirection :2719-2725).
struct Walker {
    int d1, d2, bi;
    I4 pos;
};
// One step of a corner walker: across d1 (through a connection, with the directions remapped, or a periodic wrap),
// then it turns; returns false at a FIXED face, leaving the walker where it is (getCornerValue :2775-2845).
template <int D>
__device__ __forceinline__ bool walk_step(const pict_block_desc *blocks, Walker &w) {
    const pict_block_desc *wb = blocks + w.bi;
    const int axis = w.d1 >> 1;
    if (at_bound(w.pos, w.d1, wb)) {
        const pict_boundary_desc *bd = &wb->bounds[w.d1];
        if (bd->type == PICT_BOUNDARY_FIXED) return false;
        if (bd->type == PICT_BOUNDARY_CONNECTED) {
            w.bi = bd->connected_block;
            w.pos.a[axis] += face_sign(w.d1);
            w.pos = connected_pos<D>(blocks, w.pos, axis, bd, 1);
            const int o1 = w.d1;
            w.d1 = connected_dir<D>(w.d2, axis, bd);
            w.d2 = connected_dir<D>(o1, axis, bd) ^ 1;
            return true;
        }
        if (bd->type != PICT_BOUNDARY_PERIODIC) return true;  // no other kind of face moves the walker
        w.pos.a[axis] = (w.d1 & 1) ? 0 : wb->size[axis] - 1;
    } else {
        w.pos.a[axis] += face_sign(w.d1);
    }
    const int o1 = w.d1;
    w.d1 = w.d2;
    w.d2 = o1 ^ 1;
    return true;
}

template <typename R>
struct Corner {
    R data;
    int num;    // cells averaged; 0: the value came from a FIXED boundary
    int btype;  // that boundary's condition type
};
// getCornerValue :2742-2859
template <typename R, int D>
__device__ __forceinline__ Corner<R> corner_value(const DomainK<R> &dom, int bi, const I4 &pos, int dir1, int dir2,
                                                  bool inc0, bool inc1, int type) {
    R data = 0;
    int num = 1;
    if (inc0) data += block_data(dom, bi, pos, type);
    Walker w[2] = {{dir1, dir2, bi, pos}, {dir2, dir1, bi, pos}};
    for (int depth = 1; depth <= 2; ++depth) {
        for (int d = 0; d < 2; ++d) {
            if (!walk_step<D>(dom.blocks, w[d])) {
                const pict_block_desc *wb = dom.blocks + w[d].bi;
                Corner<R> r;
                r.num = 0;
                r.btype = fixed_type(wb, w[d].d1, type, w[d].pos.a[3]);
                if (at_bound(w[d].pos, w[d].d2, wb)) {
                    r.data = fixed_data<R>(wb, w[d].pos, w[d].d1, type);
                } else {
                    I4 np = w[d].pos;
                    np.a[w[d].d2 >> 1] += face_sign(w[d].d2);
                    r.data = (fixed_data<R>(wb, w[d].pos, w[d].d1, type) + fixed_data<R>(wb, np, w[d].d1, type)) *
                             R(0.5);
                }
                return r;
            }
            if (w[d].bi == w[d ^ 1].bi && i4_eq(w[d].pos, w[d ^ 1].pos)) {
                Corner<R> r;
                r.data = data / (R)num;
                r.num = num;
                r.btype = PICT_BC_DIRICHLET;
                return r;
            }
            if (depth > 1 || inc1) data += block_data(dom, w[d].bi, w[d].pos, type);
            ++num;
        }
    }
    Corner<R> r;
    r.data = data / (R)num;
    r.num = num;
    r.btype = PICT_BC_DIRICHLET;
    return r;
}

// Neighbour across `face` for the cross-coefficient interpolation (interpolateNonOrthoLaplaceComponents
// :1935-1963): its block and cell (w = 0) and its single-axis coefficients (unmapped axes).
template <typename R, int D>
__device__ __forceinline__ void interp_neighbor(const pict_block_desc *blocks, int bi, const I4 &pos, int face,
                                                int &nbi, I4 &pn, R *aN) {
    const pict_block_desc *b = blocks + bi;
    const int axis = face >> 1, up = face & 1;
    pn = pos;
    pn.a[3] = axis;
    nbi = bi;
    if (at_bound(pos, face, b) && b->bounds[face].type == PICT_BOUNDARY_CONNECTED) {
        nbi = b->bounds[face].connected_block;
        pn = connected_pos_ch<D>(blocks, pn, axis, &b->bounds[face]);
        lap_single_axis<R, D>(blocks + nbi, pn, axis, aN);
    } else {
        const bool per = b->bounds[face].type == PICT_BOUNDARY_PERIODIC;
        if (!up && pos.a[axis] == 0 && per) pn.a[axis] = b->size[axis] - 1;
        else if (up && pos.a[axis] == b->size[axis] - 1 && per) pn.a[axis] = 0;
        else pn.a[axis] = pos.a[axis] + face_sign(face);
        lap_single_axis<R, D>(b, pn, axis, aN);
    }
    pn.a[3] = 0;
}
// condition of interpolateNonOrthoLaplaceComponents :1934: a face gets interpolated coefficients unless the cell
// is at an end of the face's axis and the face is prescribed
__device__ __forceinline__ bool interp_face_live(const pict_block_desc *b, const I4 &pos, int face) {
    const int axis = face >> 1;
    return (0 < pos.a[axis] && pos.a[axis] < b->size[axis] - 1) || !prescribed(face, b);
}
// interpolateNonOrthoLaplaceComponents :1911-1986 (cell metrics): out[face*(D-1) + i-1] is the face value of
// alpha_{axis,(axis+i)%D}, times 1/A and the viscosity when asked.
// getLaplaceCoefficientsSingleFace :1320-1343: det Minv[c].Minv[axis] of the face's own transform (the unit
// vector along the face's axis without a transform; a transformed block needs face transforms)
template <typename R, int D>
__device__ __forceinline__ R lap_single_face(const pict_block_desc *b, const I4 &pos, int face, int c) {
    const int axis = face >> 1;
    if (!b->has_transform) return c == axis ? R(1) : R(0);
    const R *T = face_T<R, D>(b, pos, face);
    R s = 0;
#pragma unroll
    for (int j = 0; j < D; ++j) s += Tminv<R, D>(T, c, j) * Tminv<R, D>(T, axis, j);
    return Tdet<R, D>(T) * s;
}

// FACE (useFaceTransform, :1962-1972): each face's cross coefficients from the face's transform times the mean of
// the two 1/A, instead of the mean of the two cells' products
template <typename R, int D, bool FACE = false>
__device__ __forceinline__ void interp_nonortho(const DomainK<R> &dom, int bi, const I4 &pos, bool with_visc,
                                                bool with_a, R *out) {
    const pict_block_desc *b = dom.blocks + bi;
    R aP[D * (D + 1) / 2];
    lap_full<R, D>(b, pos, aP);
    const R one = 1;
    const R raP = (with_a ? one / dom.A[global_index(pos, b)] : one) * (with_visc ? viscosity(dom, b, pos) : one);
#pragma unroll
    for (int face = 0; face < 2 * D; ++face) {
        const int axis = face >> 1;
        if (interp_face_live(b, pos, face)) {
            int nbi;
            I4 pn;
            R aN[D];
            interp_neighbor<R, D>(dom.blocks, bi, pos, face, nbi, pn, aN);
            const pict_block_desc *nb = dom.blocks + nbi;
            const R raN = (with_a ? one / dom.A[global_index(pn, nb)] : one) * (with_visc ? viscosity(dom, nb, pn) : one);
#pragma unroll
            for (int i = 1; i < D; ++i) {
                const int oa = (axis + i) % D;
                out[face * (D - 1) + i - 1] = FACE ? lap_single_face<R, D>(b, pos, face, oa) * (raP + raN) * R(0.5)
                                                   : (aP[lap_ci(axis, oa)] * raP + aN[oa] * raN) * R(0.5);
            }
        } else {
#pragma unroll
            for (int i = 1; i < D; ++i) out[face * (D - 1) + i - 1] = 0;
        }
    }
}

// Non-orthogonal matrix entries of one face of a row (PISO_build_matrix :3738-3800 with sign -1,
// PISO_build_pressure_matrix :4877-4932 with sign +1); rv in slot order (0 = diagonal, face+1 = neighbour).
template <typename R, int D>
__device__ __forceinline__ void nonortho_matrix_face(const DomainK<R> &dom, int bi, const I4 &pos, int face,
                                                     const R *ai, int flags, R sign, int corner_type,
                                                     bool grad_always, R *rv) {
    const int dim = face >> 1, fs = face_sign(face);
    const bool nbrs = flags & PICT_NON_ORTHO_DIRECT_MATRIX, ctr = flags & PICT_NON_ORTHO_CENTER_MATRIX;
#pragma unroll
    for (int i = 1; i < D; ++i) {
        const int ta = (dim + i) % D;
        const R alpha = ai[face * (D - 1) + i - 1];
        if (alpha == R(0)) continue;
#pragma unroll
        for (int tu = 0; tu < 2; ++tu) {
            const int tf = (ta << 1) | tu, tfs = face_sign(tf);
            const Corner<R> cv = corner_value<R, D>(dom, bi, pos, face, tf, false, false, corner_type);
            if (cv.num < 1) {
                if (grad_always || cv.btype == PICT_BC_NEUMANN) {
                    const R v = (R)(fs * tfs) * alpha * R(0.25);
                    if (ctr) rv[0] += sign * 3 * v;
                    if (nbrs) {
                        rv[face + 1] += sign * 3 * v;
                        rv[(tf ^ 1) + 1] -= sign * v;
                    }
                }
            } else {
                const R norm = (R)(1.0 / (double)cv.num);
                const R v = (R)(fs * tfs) * alpha * norm;
                if (ctr) rv[0] += sign * v;
                if (nbrs) {
                    rv[face + 1] += sign * v;
                    rv[tf + 1] += sign * v;
                }
            }
        }
    }
}

// getNonOrthoLaplaceRHS_v2 :3033-3190 (cell metrics): the explicit non-orthogonal part of the Laplacian of `type`
// at `pos` (w = component) from the previous iterate.  FACE (useFaceTransform, :3066-3068, :3128-3146): the face
// coefficients from the face's transform, times the mean of the two 1/A.
template <typename R, int D, bool FACE = false>
__device__ __forceinline__ R nonortho_rhs(const DomainK<R> &dom, int bi, const I4 &pos, int flags, int type,
                                          bool with_visc, bool with_a) {
    if (!(flags & kNonOrthoRhs)) return R(0);
    const pict_block_desc *b = dom.blocks + bi;
    const bool direct = flags & PICT_NON_ORTHO_DIRECT_RHS;
    R S = 0;
    R aP[D * (D + 1) / 2];
    lap_full<R, D>(b, pos, aP);
    const R one = 1;
    const R raP = (with_a ? one / dom.A[global_index(pos, b)] : one) * (with_visc ? viscosity(dom, b, pos) : one);
    for (int face = 0; face < 2 * D; ++face) {
        const int axis = face >> 1, fs = face_sign(face);
        if (at_bound(pos, face, b) && prescribed(face, b)) {
            if (fixed_type(b, face, type, pos.a[3]) == PICT_BC_NEUMANN || type == GD_PRESSURE_RESULT) continue;
            R bA[D];
            if constexpr (FACE) {
#pragma unroll
                for (int c = 0; c < D; ++c) bA[c] = lap_single_face<R, D>(b, pos, face, c);
            } else {
                lap_neighbour_boundary<R, D>(&b->bounds[face], pos, axis, bA);
            }
#pragma unroll
            for (int i = 1; i < D; ++i) {
                const int ta = (axis + i) % D;
                const bool tlo = pos.a[ta] == 0, thi = pos.a[ta] == b->size[ta] - 1;
                I4 lo = pos, hi = pos;
                if (!tlo) lo.a[ta] -= 1;
                if (!thi) hi.a[ta] += 1;
                const R df = (tlo || thi) ? R(1) : R(0.5);
                const R tdg = df * (fixed_data<R>(b, hi, face, type) - fixed_data<R>(b, lo, face, type));
                const R bv = with_visc ? viscosity(dom, b, pos) : one;
                S -= (R)fs * bA[ta] * tdg * bv;
            }
            continue;
        }
        const NCell n = resolve_neighbor<D>(dom.blocks, bi, pos, face);
        R aN[D];
        lap_single_axis_neighbor<R, D>(dom.blocks, n, face, aN);
        const pict_block_desc *nb = dom.blocks + n.bi;
        const R raN = (with_a ? one / dom.A[global_index(n.pos, nb)] : one) * (with_visc ? viscosity(dom, nb, n.pos) : one);
#pragma unroll
        for (int i = 1; i < D; ++i) {
            const int ta = (axis + i) % D;
            const R fa = FACE ? lap_single_face<R, D>(b, pos, face, ta) * (raP + raN) * R(0.5)
                              : (aP[lap_ci(axis, ta)] * raP + aN[ta] * raN) * R(0.5);
            R tdg = 0;
#pragma unroll
            for (int tu = 0; tu < 2; ++tu) {
                const int tf = (ta << 1) | tu, tfs = face_sign(tf);
                const Corner<R> cv = corner_value<R, D>(dom, bi, pos, face, tf, false, direct, type);
                if (cv.num == 0 && (cv.btype == PICT_BC_NEUMANN || type == GD_PRESSURE_RESULT)) {
                    const int to = tf ^ 1, tofs = -tfs;
                    if (direct) {
                        tdg += (R)tfs * data_neighbor<R, D>(dom, bi, pos, face, type) * R(0.75);
                        tdg += (R)tofs * data_neighbor<R, D>(dom, bi, pos, to, type) * R(0.25);
                    }
                    if (flags & PICT_NON_ORTHO_DIAGONAL_RHS)
                        tdg += (R)tofs * data_neighbor_diag<R, D>(dom, bi, pos, face, to, type) * R(0.25);
                } else {
                    tdg += (R)tfs * cv.data;
                }
            }
            S -= (R)fs * fa * tdg;
        }
    }
    return S;
}

}  // namespace pict
