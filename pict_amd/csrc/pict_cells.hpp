// pict_cells.hpp -- per-cell bodies of the forward PISO stage kernels.
//
// Each body handles one 256-cell tile of one block (tile = index into the domain's
// tile tables, one thread per cell).  The single-domain kernels (pict_stages.hip)
// run tile blockIdx.x of their domain; the batched kernels (pict_batch.hip) run the
// tiles of many domains in one launch, each workgroup reading its domain's view
// from a table.
#pragma once

#include <hip/hip_runtime.h>

#include "pict_device.hpp"
#include "pict_nonortho.hpp"

namespace pict {

// ---------------------------------------------------------------------------
// PISO_build_matrix :3601-3865 (velocity matrix C and its diagonal A).  Values
// only; the pattern is written by k_csr_pattern.  NO: the non-orthogonal matrix
// terms of `flags` are included (instances without them carry none of that code).
// SC: the matrix of passive-scalar channel ch (forPassiveScalar): the scalar
// viscosity, the channel's boundary types, cross coefficients times the scalar
// viscosity (:3683, :3742-3755, :3815-3826).
// ---------------------------------------------------------------------------
template <typename R, int D, bool NO = false, bool SC = false>
__device__ __forceinline__ void adv_matrix_cell(const DomainK<R> &dom, int tile, R dt, int flags = 0, int ch = 0) {
    const pict_block_desc *b;
    int bi, f;
    if (!tile_cell(dom, tile, b, bi, f)) return;
    const I4 pos = unflat(f, b);
    R fl[2 * D];
    fluxes_nd<R, D>(dom, b, pos, fl, nullptr);
    int idx[7];
    R rv[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) { idx[k] = -1; rv[k] = 0; }
    const R det = cell_det<R, D>(b, pos);
    R diag = det / dt;
    idx[0] = f + b->global_offset;
    const R visc_p = SC ? scalar_viscosity(dom, ch) : viscosity(dom, b, pos);
    R alpha_p[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        I4 t = pos;
        t.a[3] = d;
        alpha_p[d] = laplace_orth<R, D>(b, t);
    }
#pragma unroll
    for (int face = 0; face < 2 * D; ++face) {
        const int dim = face >> 1;
        const R fs = (face & 1) ? R(1) : R(-1);
        if (!(at_bound(pos, face, b) && prescribed(face, b))) {
            const R ff = fs * R(0.5) * fl[face];
            diag += ff;
            rv[face + 1] += ff;
            int nb;
            I4 t = neighbor_cell<D>(dom.blocks, b, pos, face, bi, nb);
            const pict_block_desc *nbp = dom.blocks + nb;
            const R alpha = laplace_orth<R, D>(nbp, t);
            t.a[3] = 0;
            const R visc_n = SC ? visc_p : viscosity(dom, nbp, t);
            idx[face + 1] = global_index(t, nbp);
            const R vc = (alpha_p[dim] * visc_p + alpha * visc_n) * R(0.5);
            diag += vc;
            rv[face + 1] -= vc;
            if constexpr (NO && D > 1) {
                if (flags & kNonOrthoMatrix) {
                    R ai[2 * D * (D - 1)];
                    interp_nonortho<R, D>(dom, bi, pos, !SC, false, ai);
                    I4 cp = pos;
                    if constexpr (SC) {
#pragma unroll
                        for (int k = 0; k < 2 * D * (D - 1); ++k) ai[k] *= visc_p;
                        cp.a[3] = ch;
                    }
                    rv[0] = diag;
                    nonortho_matrix_face<R, D>(dom, bi, cp, face, ai, flags, R(-1), SC ? GD_SCALAR : GD_VELOCITY,
                                               false, rv);
                    diag = rv[0];
                }
            }
        } else {
            const int bc = SC ? fixed_scalar_bc(&b->bounds[face], ch) : b->bounds[face].velocity_bc;
            const int slip = bc == PICT_BC_DIRICHLET ? 0 : 1;
            diag += R((1 - slip) * 2) * visc_p * alpha_p[dim];
        }
    }
    rv[0] = diag;
    const int g = f + b->global_offset;
    const int start = dom.csr_row[g], end = dom.csr_row[g + 1];
    for (int i = 0; i < end - start; ++i) {
        const int s = lowest_slot(idx);
        if (s < 0) break;
        rv[s] /= det;
        dom.C[start + i] = rv[s];
        idx[s] = -1;
    }
    dom.A[g] = rv[0];
}

// kPISO_build_advection_RHS :4281-4385 (NO: with the explicit non-orthogonal terms of `flags`, from the
// velocity result)
template <typename R, int D, bool NO = false>
__device__ __forceinline__ void adv_rhs_cell(const DomainK<R> &dom, int tile, R dt, int apply_pg, int flags = 0) {
    const pict_block_desc *b;
    int bi, f;
    if (!tile_cell(dom, tile, b, bi, f)) return;
    const I4 pos = unflat(f, b);
    const R det = cell_det<R, D>(b, pos);
    R pg[D];
    if (apply_pg) pressure_gradient<R, D>(dom, b, pos, pg);
    for (int d = 0; d < D; ++d) {
        I4 t = pos;
        t.a[3] = d;
        R r = det * reinterpret_cast<const R *>(b->velocity)[flat4(t, b->stride)] / dt;
        r += fixed_boundary_source<R, D>(dom, b, pos, d);
        if constexpr (NO && D > 1) r -= nonortho_rhs<R, D>(dom, bi, t, flags, GD_VELOCITY_RESULT, true, false);
        r /= det;
        r += velocity_source<R>(b, t);
        if (apply_pg) r -= pg[d] * dt;
        dom.vrhs[(size_t)d * dom.N + b->global_offset + f] = r;
    }
}

// kPISO_build_scalar_advection_RHS :4080-4190: scalar_rhs of every channel (NO: with the explicit non-orthogonal
// terms of `flags`, from the scalar result, times the scalar viscosity)
template <typename R, int D, bool NO = false>
__device__ __forceinline__ void scalar_rhs_cell(const DomainK<R> &dom, int tile, R dt, int flags = 0) {
    const pict_block_desc *b;
    int bi, f;
    if (!tile_cell(dom, tile, b, bi, f)) return;
    const I4 pos = unflat(f, b);
    const R det = cell_det<R, D>(b, pos);
    for (int c = 0; c < dom.channels; ++c) {
        I4 t = pos;
        t.a[3] = c;
        const R visc = scalar_viscosity(dom, c);
        R r = det * reinterpret_cast<const R *>(b->scalar)[flat4(t, b->stride)] / dt;
#pragma unroll
        for (int face = 0; face < 2 * D; ++face) {
            if (!at_bound(pos, face, b)) continue;
            const pict_boundary_desc *fb = &b->bounds[face];
            if (fb->type != PICT_BOUNDARY_FIXED) continue;
            const int ax = face >> 1;
            const R normal = (face & 1) ? R(1) : R(-1);
            I4 bp = pos;
            bp.a[ax] = 0;
            const R sv = fixed_scalar<R>(fb, bp, c);  // value (Dirichlet) or gradient (Neumann)
            bp.a[3] = ax;
            const R alpha = laplace_orth_fixed<R, D>(fb, bp);
            const R flux = fixed_contravariant<R, D>(fb, bp) * normal;
            r -= sv * flux;
            r += fixed_scalar_bc(fb, c) == PICT_BC_DIRICHLET ? sv * visc * R(2) * alpha : sv * visc;
        }
        if constexpr (NO && D > 1) r -= nonortho_rhs<R, D>(dom, bi, t, flags, GD_SCALAR_RESULT, false, false) * visc;
        r /= det;
        dom.srhs[(size_t)c * dom.N + b->global_offset + f] = r;
    }
}

// PISO_build_pressure_matrix :4797-4963 (cell metrics; NO: with the non-orthogonal matrix terms of `flags`; FACE:
// the orthogonal coefficient from the face's transform times the mean of the two 1/A, useFaceTransform :4857-4865)
template <typename R, int D, bool NO = false, bool FACE = false>
__device__ __forceinline__ void p_matrix_cell(const DomainK<R> &dom, int tile, int flags = 0) {
    const pict_block_desc *b;
    int bi, f;
    if (!tile_cell(dom, tile, b, bi, f)) return;
    const I4 pos = unflat(f, b);
    const int g = b->global_offset + f;
    int idx[7];
    R rv[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) { idx[k] = -1; rv[k] = 0; }
    R alphaP[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        I4 t = pos;
        t.a[3] = d;
        alphaP[d] = laplace_orth<R, D>(b, t);
    }
    const R raP = R(1.0) / dom.A[g];
    idx[0] = g;
#pragma unroll
    for (int face = 0; face < 2 * D; ++face) {
        if (at_bound(pos, face, b) && prescribed(face, b)) continue;
        const int dim = face >> 1;
        int nb;
        I4 t = neighbor_cell<D>(dom.blocks, b, pos, face, bi, nb);
        const pict_block_desc *nbp = dom.blocks + nb;
        const R alphaN = laplace_orth<R, D>(nbp, t);
        t.a[3] = 0;
        const int gn = global_index(t, nbp);
        const R raN = R(1.0) / dom.A[gn];
        const R coef = FACE ? laplace_orth_face<R, D>(b, pos, face) * (R(0.5) * (raP + raN))
                            : R(0.5) * (alphaP[dim] * raP + alphaN * raN);
        rv[0] -= coef;
        rv[face + 1] += coef;
        idx[face + 1] = gn;
        if constexpr (NO && D > 1) {
            if (flags & kNonOrthoMatrix) {
                R ai[2 * D * (D - 1)];
                interp_nonortho<R, D, FACE>(dom, bi, pos, false, true, ai);
                nonortho_matrix_face<R, D>(dom, bi, pos, face, ai, flags, R(1), GD_IS_FIXED, true, rv);
            }
        }
    }
    const int start = dom.csr_row[g], end = dom.csr_row[g + 1];
    for (int i = 0; i < end - start; ++i) {
        const int s = lowest_slot(idx);
        if (s < 0) break;
        dom.P[start + i] = rv[s];
        idx[s] = -1;
    }
}

// PISO_build_pressure_rhs :5121-5240
template <typename R, int D>
__device__ __forceinline__ void p_rhs_cell(const DomainK<R> &dom, int tile, R dt) {
    const pict_block_desc *b;
    int bi, f;
    if (!tile_cell(dom, tile, b, bi, f)) return;
    const I4 pos = unflat(f, b);
    const int g = b->global_offset + f;
    const R rDiag = R(1) / dom.A[g];
    const R det = cell_det<R, D>(b, pos);
    const int cs = dom.csr_row[g], ce = dom.csr_row[g + 1];
    int col[7];
    R val[7];
    const int rs = ce - cs < 7 ? ce - cs : 7;
    for (int k = 0; k < rs; ++k) {
        col[k] = dom.csr_index[cs + k];
        val[k] = dom.C[cs + k];
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
        I4 t = pos;
        t.a[3] = d;
        const R velOld = reinterpret_cast<const R *>(b->velocity)[flat4(t, b->stride)] / dt;
        R H = 0;
        for (int k = 0; k < rs; ++k)
            if (col[k] != g) H += val[k] * dom.vres[col[k] + (size_t)dom.N * d];
        R S = fixed_boundary_source<R, D>(dom, b, pos, d);
        S /= det;
        S += velocity_source<R>(b, t);
        dom.prhs[g + (size_t)dom.N * d] = rDiag * (velOld - H + S);
    }
}

// k_computePressureRHSdivergenceFromFlux :5374-5419 (vg = pressure RHS) and
// k_computeVelocityDivergenceFromFlux :6333-6354 (vg = nullptr, block velocity).
template <typename R, int D>
__device__ __forceinline__ void divergence_cell(const DomainK<R> &dom, int tile, const R *vg, R *out, R dt, int tsn) {
    const pict_block_desc *b;
    int bi, f;
    if (!tile_cell(dom, tile, b, bi, f)) return;
    const I4 pos = unflat(f, b);
    R fl[2 * D];
    fluxes_nd<R, D>(dom, b, pos, fl, vg);
    R div = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) div += fl[2 * d + 1] - fl[2 * d];
    if (tsn) div /= dt;
    out[f + b->global_offset] = div;
}

// slerp2Halfpoint :937-971 with eps 1e-5: the direction of the mean of the two vectors (both scaled by the first one's
// inverse length, as the reference does: the direction is that of v1 + v2, the smallness test sees the scale) with
// the mean of their lengths; the plain mean when a vector or the direction is too short
template <typename R, int D>
__device__ __forceinline__ void slerp2_half(const R (&v1)[D], const R (&v2)[D], R (&out)[D]) {
    const R eps = R(1e-5);
    R s1 = 0, s2 = 0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        s1 += v1[k] * v1[k];
        s2 += v2[k] * v2[k];
    }
    R n[D], sn = 0;
    const R r1 = R(1) / sqrt(s1);
#pragma unroll
    for (int k = 0; k < D; ++k) {
        n[k] = (v1[k] * r1 + v2[k] * r1) * R(0.5);
        sn += n[k] * n[k];
    }
    if (s1 < eps || s2 < eps || sn < eps) {
#pragma unroll
        for (int k = 0; k < D; ++k) out[k] = (v1[k] + v2[k]) * R(0.5);
        return;
    }
    const R rn = R(1) / sqrt(sn);
    const R m = (R)(((double)sqrt(s1) + (double)sqrt(s2)) * 0.5);
#pragma unroll
    for (int k = 0; k < D; ++k) out[k] = n[k] * rn * m;
}

// computeFluxesWithFaceTransforms :1703-1785: per face, the neighbour's vector of vg (a fixed face: the boundary
// velocity), blended with the cell's by slerp2_half (at fixed faces too, as in the reference), projected on the
// face's det * Minv[axis]; the divergence as k_computePressureRHSdivergenceFromFlux :5374-5419
template <typename R, int D>
__device__ __forceinline__ void divergence_face_cell(const DomainK<R> &dom, int tile, const R *vg, R *out, R dt,
                                                     int tsn) {
    const pict_block_desc *b;
    int bi, f;
    if (!tile_cell(dom, tile, b, bi, f)) return;
    const I4 pos = unflat(f, b);
    const int g = b->global_offset + f;
    R vc[D];
#pragma unroll
    for (int k = 0; k < D; ++k) vc[k] = vg[g + (size_t)dom.N * k];
    R div = 0;
#pragma unroll
    for (int face = 0; face < 2 * D; ++face) {
        const int dim = face >> 1, up = face & 1;
        R vn[D];
        I4 t = pos;
        t.a[3] = dim;
        int gn = -1;
        if (at_bound(pos, face, b)) {
            const pict_boundary_desc *fb = &b->bounds[face];
            if (fb->type == PICT_BOUNDARY_FIXED) {
                I4 bp = pos;
                bp.a[dim] = 0;
#pragma unroll
                for (int k = 0; k < D; ++k) vn[k] = fixed_vel<R>(fb, bp, k);
            } else if (fb->type == PICT_BOUNDARY_CONNECTED) {
                const I4 o = connected_pos<D>(dom.blocks, t, dim, fb, 0);
                gn = global_index(o, dom.blocks + fb->connected_block);
            } else {
                t.a[dim] = up ? 0 : b->size[dim] - 1;
                gn = global_index(t, b);
            }
        } else {
            t.a[dim] += up ? 1 : -1;
            gn = global_index(t, b);
        }
        if (gn >= 0)
#pragma unroll
            for (int k = 0; k < D; ++k) vn[k] = vg[gn + (size_t)dom.N * k];
        R v[D];
        slerp2_half<R, D>(vc, vn, v);
        const R *F = face_T<R, D>(b, pos, face);
        R s = 0;
#pragma unroll
        for (int j = 0; j < D; ++j) s += Tminv<R, D>(F, dim, j) * v[j];
        const R flux = Tdet<R, D>(F) * s;
        div += up ? flux : -flux;
    }
    if (tsn) div /= dt;
    out[g] = div;
}

// k_pressureRHSaddNonOrthoComponents :5455-5477: pressure_rhs_div += the explicit non-orthogonal part of the
// pressure Laplacian from the pressure result
template <typename R, int D, bool FACE = false>
__device__ __forceinline__ void p_rhs_nonortho_cell(const DomainK<R> &dom, int tile, int flags) {
    const pict_block_desc *b;
    int bi, f;
    if (!tile_cell(dom, tile, b, bi, f)) return;
    const int g = f + b->global_offset;
    R S = dom.prhs_div[g];
    if constexpr (D > 1)
        S += nonortho_rhs<R, D, FACE>(dom, bi, unflat(f, b), flags, GD_PRESSURE_RESULT, false, true);
    dom.prhs_div[g] = S;
}

// PISO_update_velocity :5947-5980
template <typename R, int D>
__device__ __forceinline__ void correct_velocity_cell(const DomainK<R> &dom, int tile, R dt, int tsn) {
    const pict_block_desc *b;
    int bi, f;
    if (!tile_cell(dom, tile, b, bi, f)) return;
    const I4 pos = unflat(f, b);
    const int g = b->global_offset + f;
    const R rDiag = R(1) / dom.A[g];
    R pg[D];
    pressure_gradient<R, D>(dom, b, pos, pg);
#pragma unroll
    for (int d = 0; d < D; ++d) {
        R u = -rDiag * pg[d];
        if (tsn) u *= dt;
        u += dom.prhs[g + (size_t)dom.N * d];
        dom.vres[g + (size_t)dom.N * d] = u;
    }
}

// getPressureGradientFVM :991-1124 (MODE 0: mean of the two cells' fluxes det * Minv[dim] * p; MODE 3: the face's
// metrics times the mean pressure).  A fixed face extrapolates the pressure from the cell and its inner neighbour:
// half a cell, with the boundary's transform, for MODE 0; a whole cell, then averaged with the cell, for MODE 3.
// Across a connection the neighbour contributes its own axis' row, negated when both faces are on the same side.
// Without a transform: the finite-difference gradient.
template <typename R, int D, int MODE>
__device__ __forceinline__ void pressure_gradient_fvm(const DomainK<R> &dom, const pict_block_desc *b, I4 pos,
                                                      R *g) {
    if (!b->has_transform) {
        pressure_gradient<R, D>(dom, b, pos, g);
        return;
    }
    pos.a[3] = 0;
    const R *TP = cell_T<R, D>(b, pos);
    const R pP = reinterpret_cast<const R *>(b->pressure)[flat4(pos, b->stride)];
    const R detP = Tdet<R, D>(TP);
    R acc[D];
#pragma unroll
    for (int j = 0; j < D; ++j) acc[j] = 0;
#pragma unroll
    for (int face = 0; face < 2 * D; ++face) {
        const int dim = face >> 1, up = face & 1;
        const R sgn = up ? R(1) : R(-1);
        const R *T;
        int axn = dim;
        R tsign = 1, p;
        I4 t = pos;
        if (at_bound(pos, face, b)) {
            const pict_boundary_desc *fb = &b->bounds[face];
            if (fb->type == PICT_BOUNDARY_FIXED) {
                t.a[dim] = 0;
                T = bound_T<R, D>(fb, t);
                t.a[dim] = pos.a[dim] - (up ? 1 : -1);
                const R pin = reinterpret_cast<const R *>(b->pressure)[flat4(t, b->stride)];
                if (MODE != 3) {
                    p = pP + (pP - pin) * R(0.5);
                    const R s = Tdet<R, D>(T) * p * sgn;
#pragma unroll
                    for (int j = 0; j < D; ++j) acc[j] += Tminv<R, D>(T, dim, j) * s;
                    continue;
                }
                p = pP + (pP - pin);
            } else if (fb->type == PICT_BOUNDARY_CONNECTED) {
                const pict_block_desc *ob = dom.blocks + fb->connected_block;
                t.a[3] = dim;
                I4 o = connected_pos_ch<D>(dom.blocks, t, dim, fb);
                axn = o.a[3];
                if ((fb->axes[0] & 1) == up) tsign = -1;
                o.a[3] = 0;
                T = cell_T<R, D>(ob, o);
                p = reinterpret_cast<const R *>(ob->pressure)[flat4(o, ob->stride)];
            } else {  // periodic
                t.a[dim] = up ? 0 : b->size[dim] - 1;
                T = cell_T<R, D>(b, t);
                p = reinterpret_cast<const R *>(b->pressure)[flat4(t, b->stride)];
            }
        } else {
            t.a[dim] += up ? 1 : -1;
            T = cell_T<R, D>(b, t);
            p = reinterpret_cast<const R *>(b->pressure)[flat4(t, b->stride)];
        }
        if (MODE == 3) {
            const R *F = face_T<R, D>(b, pos, face);
            const R s = Tdet<R, D>(F) * (p + pP) * R(0.5) * sgn;
#pragma unroll
            for (int j = 0; j < D; ++j) acc[j] += Tminv<R, D>(F, dim, j) * s;
        } else {
            const R sP = detP * pP, sN = Tdet<R, D>(T) * p * tsign, h = R(0.5) * sgn;
#pragma unroll
            for (int j = 0; j < D; ++j) acc[j] += (Tminv<R, D>(TP, dim, j) * sP + Tminv<R, D>(T, axn, j) * sN) * h;
        }
    }
    const R rdet = R(1) / detP;
#pragma unroll
    for (int j = 0; j < D; ++j) g[j] = acc[j] * rdet;
}

// PISO_update_velocity_PressureFVM :6013-6049 (CorrectVelocity versions 5 and 6)
template <typename R, int D, int MODE>
__device__ __forceinline__ void correct_velocity_fvm_cell(const DomainK<R> &dom, int tile, R dt, int tsn) {
    const pict_block_desc *b;
    int bi, f;
    if (!tile_cell(dom, tile, b, bi, f)) return;
    const I4 pos = unflat(f, b);
    const int g = b->global_offset + f;
    const R rDiag = R(1) / dom.A[g];
    R pg[D];
    pressure_gradient_fvm<R, D, MODE>(dom, b, pos, pg);
#pragma unroll
    for (int d = 0; d < D; ++d) {
        R u = -rDiag * pg[d];
        if (tsn) u *= dt;
        dom.vres[g + (size_t)dom.N * d] = u + dom.prhs[g + (size_t)dom.N * d];
    }
}

// getTransformMetricOrthogonal :1181-1192: Minv[w][w] of a cell (1 without a transform)
template <typename R, int D>
__device__ __forceinline__ R metric_orth(const pict_block_desc *b, const I4 &p) {
    if (!b->has_transform) return R(1);
    return Tminv<R, D>(cell_T<R, D>(b, p), p.a[3], p.a[3]);
}
template <typename R, int D>
__device__ __forceinline__ R det_or_one(const pict_block_desc *b, const I4 &p) {
    return b->has_transform ? Tdet<R, D>(cell_T<R, D>(b, p)) : R(1);
}

// PISO_update_velocity_v4_orthogonal :6052-6160 (CorrectVelocity version 4): per component a face flux from the
// two cells' covariant H/A and pressure difference, a quarter per side, over det*t.  The cell's own metric t is read
// at channel 0 for every component (the reference's unflattened cell index carries w = 0), the neighbours' at the
// component's channel, remapped across connections.  Prescribed faces add nothing; the time step always scales the
// pressure term.
template <typename R, int D>
__device__ __forceinline__ void correct_velocity_v4_cell(const DomainK<R> &dom, int tile, R dt) {
    const pict_block_desc *b;
    int bi, f;
    if (!tile_cell(dom, tile, b, bi, f)) return;
    const I4 pos = unflat(f, b);
    const int g = b->global_offset + f;
    const R ra = R(1) / dom.A[g];
    const R p = reinterpret_cast<const R *>(b->pressure)[f];
    const R det = det_or_one<R, D>(b, pos);
    const R t = metric_orth<R, D>(b, pos);
    const R cc = det * t;
#pragma unroll
    for (int dim = 0; dim < D; ++dim) {
        const size_t gc = g + (size_t)dom.N * dim;
        const R hbyA = dom.prhs[gc];
        R vel = 0;
#pragma unroll
        for (int up = 0; up < 2; ++up) {
            const int face = 2 * dim + up;
            const bool ab = at_bound(pos, face, b);
            if (ab && prescribed(face, b)) continue;
            const pict_block_desc *nb = b;
            I4 q = pos;
            q.a[3] = dim;
            if (ab && b->bounds[face].type == PICT_BOUNDARY_CONNECTED) {
                nb = dom.blocks + b->bounds[face].connected_block;
                q = connected_pos_ch<D>(dom.blocks, q, dim, &b->bounds[face]);
            } else if (ab) {
                q.a[dim] = up ? 0 : b->size[dim] - 1;
            } else {
                q.a[dim] += up ? 1 : -1;
            }
            const int gn = global_index(q, nb);
            const R hbyAN = dom.prhs[gn + (size_t)dom.N * q.a[3]];
            const R tN = metric_orth<R, D>(nb, q);
            q.a[3] = 0;
            const R detN = det_or_one<R, D>(nb, q);
            const R pN = reinterpret_cast<const R *>(nb->pressure)[flat4(q, nb->stride)];
            const R raN = R(1) / dom.A[gn];
            const R ccN = detN * tN;
            const R dp = up ? pN - p : p - pN;
            vel += (cc * hbyA + ccN * hbyAN - (cc * t * ra + ccN * tN * raN) * dp * dt) * R(0.25);
        }
        dom.vres[gc] = vel / (det * t);
    }
}

// Copy*Result*Blocks :6541-6733 and Copy*ResultGrad*Blocks :6819-6915 -- one
// launch instead of a memcpy per block and component.  `field` selects the
// block tensor: 0 velocity, 1 pressure, 2 velocity_grad, 3 pressure_grad,
// 4 passive scalar, 5 passive-scalar grad; `gv` is the matching global vector
// (component- or channel-major).
template <typename R>
__device__ __forceinline__ void copy_result_cell(const DomainK<R> &dom, int tile, R *gv, int field, int to_blocks) {
    const pict_block_desc *b;
    int bi, f;
    if (!tile_cell(dom, tile, b, bi, f)) return;
    const int g = b->global_offset + f;
    void *bp = field == 0 ? b->velocity : field == 1 ? b->pressure : field == 2 ? b->velocity_grad
             : field == 3 ? b->pressure_grad : field == 4 ? b->scalar : b->scalar_grad;
    R *bv = reinterpret_cast<R *>(bp);
    const int comps = field >= 4 ? dom.channels : (field & 1) ? 1 : dom.D;
    for (int d = 0; d < comps; ++d) {
        if (to_blocks) bv[(size_t)d * b->stride[3] + f] = gv[(size_t)d * dom.N + g];
        else gv[(size_t)d * dom.N + g] = bv[(size_t)d * b->stride[3] + f];
    }
}

// solve start vector x0 = 0 (field 0: velocity result, 1: pressure result) on one tile
template <typename R>
__device__ __forceinline__ void zero_result_cell(const DomainK<R> &dom, int tile, int field) {
    const pict_block_desc *b;
    int bi, f;
    if (!tile_cell(dom, tile, b, bi, f)) return;
    const int g = b->global_offset + f;
    if (field == 0) {
        for (int d = 0; d < dom.D; ++d) dom.vres[(size_t)d * dom.N + g] = R(0);
    } else {
        dom.pres[g] = R(0);
    }
}

}  // namespace pict
