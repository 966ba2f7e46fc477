// pict_grad.hip -- backward (vector-Jacobian) kernels of the PISO stages.
//
// Each kernel is the exact VJP of the forward kernel of the same stage in
// pict_stages.hip, with the reference's output contract: which gradient
// tensors are written, and whether a kernel assigns (=) or accumulates (+=).
// Scatters to cells other than the thread's own use device-scope float
// atomics (the gradient tensors are zero-initialised by the caller, as in
// the reference's Create*Grad functions).
#include <hip/hip_runtime.h>

#include "pict_host.hpp"
#include "pict_nonortho.hpp"

namespace pict {

template <typename R>
__device__ __forceinline__ void add_to(R *p, R v) {
    unsafeAtomicAdd(p, v);
}

// Per-thread partial sums of the gradient contributions whose target is the same for the
// whole workgroup (a workgroup covers cells of one block): the domain viscosity, the block's
// static viscosity and its static velocity source.  Summed over the workgroup at the end of
// the kernel and added with one atomic each, instead of one atomic per cell and face on a
// single address.
template <typename R>
struct UniformAcc {
    double visc_dom = 0.0, visc_own = 0.0, src[3] = {0.0, 0.0, 0.0};
    double visc_scalar = 0.0;  // the scalar viscosity of one passive-scalar channel
};

// workgroup sum of v (all threads call it; result valid in thread 0)
__device__ __forceinline__ double wg_sum(double v) {
    __shared__ double lds[kCellsPerTile / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
#pragma unroll
        for (int w = 0; w < kCellsPerTile / 64; ++w) t += lds[w];
    return t;
}

// Adds the workgroup's uniform-target sums (every thread of the workgroup must call it).
template <typename R, int D>
__device__ __forceinline__ void flush_uniform(const DomainK<R> &dom, const pict_block_desc *b, const UniformAcc<R> &u,
                                              bool want_src) {
    const double vd = wg_sum(u.visc_dom);
    const double vo = wg_sum(u.visc_own);
    double sv[D];
#pragma unroll
    for (int d = 0; d < D; ++d) sv[d] = want_src ? wg_sum(u.src[d]) : 0.0;
    if (threadIdx.x == 0) {
        if (dom.visc_grad && vd != 0.0) add_to(dom.visc_grad, (R)vd);
        if (b->viscosity && b->viscosity_static && b->viscosity_grad && vo != 0.0)
            add_to(reinterpret_cast<R *>(b->viscosity_grad), (R)vo);
        if (want_src && b->velocity_source_grad && b->velocity_source_static)
#pragma unroll
            for (int d = 0; d < D; ++d)
                if (sv[d] != 0.0) add_to(reinterpret_cast<R *>(b->velocity_source_grad) + d, (R)sv[d]);
    }
}

// scatterViscosity_GRAD(forPassiveScalar) :1834-1849: the workgroup's sum of scalar-viscosity gradients of channel
// ch, into the scalar viscosity gradient, or the domain viscosity gradient when there is none (every thread calls it)
template <typename R>
__device__ __forceinline__ void flush_scalar_viscosity(const DomainK<R> &dom, double v, int ch) {
    const double t = wg_sum(v);
    if (threadIdx.x == 0 && t != 0.0) {
        R *dst = dom.svisc_grad ? dom.svisc_grad + (dom.svisc_static ? 0 : ch) : dom.visc_grad;
        if (dst) add_to(dst, (R)t);
    }
}

// Position of each row slot (diag, -x, +x, ...) in the cell's sorted CSR row
// (LoadCSRrowNeighborSorted :3503-3563): rank of the slot's column among the
// row's columns.  -1 for prescribed faces.
template <typename R, int D>
__device__ __forceinline__ void slot_positions(const DomainK<R> &dom, const pict_block_desc *b, int bi, const I4 &pos,
                                               int g, int *at) {
    int idx[7];
    row_slots<R, D>(dom, b, bi, pos, idx);
    const int start = dom.csr_row[g];
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        if (idx[k] < 0) {
            at[k] = -1;
            continue;
        }
        int r = 0;
#pragma unroll
        for (int j = 0; j < 7; ++j) r += (idx[j] >= 0 && idx[j] < idx[k]) ? 1 : 0;
        at[k] = start + r;
    }
}

// scatterViscosityBlock_GRAD :1852-1873 (+ scatterViscosity_GRAD :1834-1849); `own` is the
// workgroup's block: its uniform targets go to the accumulator u.
template <typename R>
__device__ __forceinline__ void scatter_viscosity(const DomainK<R> &dom, const pict_block_desc *b, I4 p, R g,
                                                  const pict_block_desc *own, UniformAcc<R> &u) {
    R *dst;
    if (!b->viscosity) {
        u.visc_dom += (double)g;
        return;
    } else if (b->viscosity_static) {
        if (b == own) {
            u.visc_own += (double)g;
            return;
        }
        dst = reinterpret_cast<R *>(b->viscosity_grad);
    } else {
        p.a[3] = 0;
        dst = b->viscosity_grad ? reinterpret_cast<R *>(b->viscosity_grad) + flat4(p, b->stride) : nullptr;
    }
    if (dst) add_to(dst, g);
}

// scatterBlockVelocitySource_GRAD :2255-2264 (p.w = component); a static source of the
// workgroup's own block goes to the accumulator u.
template <typename R>
__device__ __forceinline__ void scatter_velocity_source(const pict_block_desc *b, const I4 &p, R g, UniformAcc<R> &u) {
    if (!b->velocity_source_grad) return;
    if (b->velocity_source_static) {
        u.src[p.a[3]] += (double)g;
        return;
    }
    add_to(reinterpret_cast<R *>(b->velocity_source_grad) + flat4(p, b->stride), g);
}

// scatterFixedBoundaryData (VELOCITY_GRAD) :2451-2476, p on the boundary grid, w = component
template <typename R>
__device__ __forceinline__ void scatter_fixed_velocity(const pict_boundary_desc *fb, const I4 &p, R g) {
    if (!fb->velocity_grad) return;
    R *dst = reinterpret_cast<R *>(fb->velocity_grad);
    add_to(fb->velocity_static ? dst + p.a[3] : dst + flat4(p, fb->stride), g);
}

// scatterContravariantComponentGradBoundaryFixed :700-716 (+ :545-560).  Like the
// reference, the transformed case scatters to all components whatever the
// boundary condition type.
template <typename R, int D>
__device__ __forceinline__ void scatter_fixed_contravariant(const pict_boundary_desc *fb, I4 p, R g) {
    if (!fb->velocity_grad) return;
    const int comp = p.a[3];
    if (fb->has_transform) {
        const R *T = bound_T<R, D>(fb, p);
        const R s = Tdet<R, D>(T) * g;
#pragma unroll
        for (int j = 0; j < D; ++j) {
            p.a[3] = j;
            scatter_fixed_velocity<R>(fb, p, Tminv<R, D>(T, comp, j) * s);
        }
    } else {
        scatter_fixed_velocity<R>(fb, p, g);
    }
}

// scatterContravariantComponentGrad :684-698 (+ VelocityGradFromContravariantComponentGrad :527-543):
// into the block velocity gradient, or into a global component-major vector `vgg`.
template <typename R, int D>
__device__ __forceinline__ void scatter_contravariant(const DomainK<R> &dom, const pict_block_desc *b, I4 p, R g,
                                                      R *vgg) {
    if (!vgg && !b->velocity_grad) return;
    const int comp = p.a[3];
    if (b->has_transform) {
        const R *T = cell_T<R, D>(b, p);
        const R s = Tdet<R, D>(T) * g;
#pragma unroll
        for (int j = 0; j < D; ++j) {
            p.a[3] = j;
            const R v = Tminv<R, D>(T, comp, j) * s;
            if (vgg) add_to(vgg + (size_t)j * dom.N + global_index(p, b), v);
            else add_to(reinterpret_cast<R *>(b->velocity_grad) + flat4(p, b->stride), v);
        }
    } else {
        if (vgg) add_to(vgg + (size_t)comp * dom.N + global_index(p, b), g);
        else add_to(reinterpret_cast<R *>(b->velocity_grad) + flat4(p, b->stride), g);
    }
}

// ScatterFluxesGradNDLoop :1635-1699: transpose of fluxes_nd.
template <typename R, int D>
__device__ __forceinline__ void scatter_fluxes(const DomainK<R> &dom, const pict_block_desc *b, const I4 &pos,
                                               const R *fg, R *vgg) {
#pragma unroll
    for (int dim = 0; dim < D; ++dim) {
        I4 t = pos;
        t.a[3] = dim;
        R own = 0;
#pragma unroll
        for (int up = 0; up < 2; ++up) {
            const int face = 2 * dim + up;
            const R g = fg[face];
            if (g == R(0)) continue;
            t.a[dim] = pos.a[dim];
            if (at_bound(pos, face, b)) {
                const pict_boundary_desc *bd = &b->bounds[face];
                if (bd->type == PICT_BOUNDARY_FIXED) {
                    t.a[dim] = 0;
                    scatter_fixed_contravariant<R, D>(bd, t, g);
                } else if (bd->type == PICT_BOUNDARY_CONNECTED) {
                    const R of = ((bd->axes[0] & 1) == up) ? R(-0.5) : R(0.5);
                    const pict_block_desc *ob = dom.blocks + bd->connected_block;
                    I4 o = connected_pos<D>(dom.blocks, t, dim, bd, 0);
                    o.a[3] = bd->axes[0] >> 1;
                    scatter_contravariant<R, D>(dom, ob, o, of * g, vgg);
                    own += R(0.5) * g;
                } else if (bd->type == PICT_BOUNDARY_PERIODIC) {
                    t.a[dim] = up ? 0 : b->size[dim] - 1;
                    scatter_contravariant<R, D>(dom, b, t, R(0.5) * g, vgg);
                    own += R(0.5) * g;
                }
            } else {
                t.a[dim] = pos.a[dim] + (up ? 1 : -1);
                scatter_contravariant<R, D>(dom, b, t, R(0.5) * g, vgg);
                own += R(0.5) * g;
            }
        }
        t.a[dim] = pos.a[dim];
        if (own != R(0)) scatter_contravariant<R, D>(dom, b, t, own, vgg);
    }
}

// VJP of fixed_boundary_source (boundary terms of the advection / pressure RHS,
// kPISO_build_advection_RHS_GRAD :4415-4465, PISO_build_pressure_rhs_GRAD :5302-5355);
// G = gradient of the source term of component d (already divided by det).
template <typename R, int D>
__device__ __forceinline__ void fixed_boundary_source_grad(const DomainK<R> &dom, const pict_block_desc *b,
                                                           const I4 &pos, int d, R G, UniformAcc<R> &u) {
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
        const R slip = fb->velocity_bc == PICT_BC_DIRICHLET ? R(0) : R(1);
        const R visc = viscosity(dom, b, pos);
        const R flux_grad = -vel * G;
        const R vel_grad = -flux * G + (R(1) - slip) * visc * R(2) * alpha * G;
        const R visc_grad = vel * (R(1) - slip) * R(2) * alpha * G;
        scatter_fixed_contravariant<R, D>(fb, t, flux_grad * normal);
        t.a[3] = d;
        scatter_fixed_velocity<R>(fb, t, vel_grad);
        scatter_viscosity<R>(dom, b, pos, visc_grad, b, u);
    }
}

// ---------------------------------------------------------------------------
// Non-orthogonal terms: VJPs of the helpers in pict_nonortho.hpp
// ---------------------------------------------------------------------------

// scatterBlockData :2340-2347 into the result gradients (p.w = component)
template <typename R>
__device__ __forceinline__ void scatter_block_data(const DomainK<R> &dom, int bi, const I4 &p, int type, R g) {
    R *dst = type == GD_VELOCITY_RESULT ? dom.vres_grad : type == GD_PRESSURE_RESULT ? dom.pres_grad
           : type == GD_SCALAR_RESULT ? dom.sres_grad : nullptr;
    if (dst) add_to(dst + (size_t)p.a[3] * dom.N + global_index(p, dom.blocks + bi), g);
}
// scatterFixedBoundaryData :2448-2478 (FixedBoundary velocity and passive-scalar gradients)
template <typename R>
__device__ __forceinline__ void scatter_fixed_data(const pict_block_desc *b, I4 p, int face, int type, R g) {
    const pict_boundary_desc *fb = &b->bounds[face];
    if (fb->type != PICT_BOUNDARY_FIXED) return;
    if (gd_scalar(type)) {
        if (!fb->scalar_grad) return;
        R *dst = reinterpret_cast<R *>(fb->scalar_grad);
        if (fb->scalar_static) {
            add_to(dst + p.a[3], g);
            return;
        }
        p.a[face >> 1] = 0;
        add_to(dst + flat4(p, fb->stride), g);
        return;
    }
    if (!(type == GD_VELOCITY_RESULT || type == GD_VELOCITY)) return;
    if (!fb->velocity_grad) return;
    R *dst = reinterpret_cast<R *>(fb->velocity_grad);
    if (fb->velocity_static) {
        add_to(dst + p.a[3], g);
        return;
    }
    p.a[face >> 1] = 0;
    add_to(dst + flat4(p, fb->stride), g);
}
// scatterBlockDataNeighbor :2560-2602 (values of FIXED faces get no gradient)
template <typename R, int D>
__device__ __forceinline__ void scatter_data_neighbor(const DomainK<R> &dom, int bi, const I4 &pos, int dir, int type,
                                                      R g) {
    int tb = bi;
    I4 t = pos;
    if (step_data<D>(dom.blocks, tb, t, dir)) scatter_block_data(dom, tb, t, type, g);
}
// scatterBlockDataNeighborDiagonal :2659-2700
template <typename R, int D>
__device__ __forceinline__ void scatter_data_neighbor_diag(const DomainK<R> &dom, int bi, const I4 &pos, int dir1,
                                                           int dir2, int type, R g) {
    const bool d1e = prescribed(dir1, dom.blocks + bi);
    int tb = bi;
    I4 t = pos;
    if (!step_data<D>(dom.blocks, tb, t, d1e ? dir2 : dir1)) return;
    if (!step_data<D>(dom.blocks, tb, t, d1e ? dir1 : dir2)) return;
    scatter_block_data(dom, tb, t, type, g);
}
// scatterCornerValue_GRAD :2869-2966: gradient g of a corner value whose forward walk averaged `num` cells
template <typename R, int D>
__device__ __forceinline__ void scatter_corner_value(const DomainK<R> &dom, int bi, const I4 &pos, int dir1, int dir2,
                                                     bool inc1, int type, R g, int num) {
    const bool from_cells = num > 0;
    R dg = g;
    if (from_cells) dg /= (R)num;
    Walker w[2] = {{dir1, dir2, bi, pos}, {dir2, dir1, bi, pos}};
    for (int depth = 1; depth <= 2; ++depth) {
        for (int d = 0; d < 2; ++d) {
            if (!walk_step<D>(dom.blocks, w[d])) {
                const pict_block_desc *wb = dom.blocks + w[d].bi;
                if (at_bound(w[d].pos, w[d].d2, wb)) {
                    scatter_fixed_data<R>(wb, w[d].pos, w[d].d1, type, dg);
                } else {
                    I4 np = w[d].pos;
                    np.a[w[d].d2 >> 1] += face_sign(w[d].d2);
                    dg *= R(0.5);
                    scatter_fixed_data<R>(wb, w[d].pos, w[d].d1, type, dg);
                    scatter_fixed_data<R>(wb, np, w[d].d1, type, dg);
                }
                return;
            }
            if (w[d].bi == w[d ^ 1].bi && i4_eq(w[d].pos, w[d ^ 1].pos)) return;
            if (from_cells && (depth > 1 || inc1)) scatter_block_data(dom, w[d].bi, w[d].pos, type, dg);
        }
    }
}

// scatterNonOrthoLaplaceComponents_GRAD :1989-2098: gradients of the interpolated cross coefficients (layout of
// interp_nonortho) to the viscosities (UniformAcc u for the workgroup's own uniform targets) and A.
template <typename R, int D>
__device__ __forceinline__ void scatter_interp_nonortho(const DomainK<R> &dom, int bi, const I4 &pos, const R *ai_grad,
                                                        bool with_visc, bool with_a, UniformAcc<R> &u) {
    const pict_block_desc *b = dom.blocks + bi;
    R aP[D * (D + 1) / 2];
    lap_full<R, D>(b, pos, aP);
    const R one = 1;
    R raP_grad = 0;
#pragma unroll
    for (int face = 0; face < 2 * D; ++face) {
        if (!interp_face_live(b, pos, face)) continue;
        const int axis = face >> 1;
        int nbi;
        I4 pn;
        R aN[D];
        interp_neighbor<R, D>(dom.blocks, bi, pos, face, nbi, pn, aN);
        const pict_block_desc *nb = dom.blocks + nbi;
        R raN_grad = 0;
#pragma unroll
        for (int i = 1; i < D; ++i) {
            const int oa = (axis + i) % D;
            const R g = ai_grad[face * (D - 1) + i - 1];
            raP_grad += aP[lap_ci(axis, oa)] * R(0.5) * g;
            raN_grad += aN[oa] * R(0.5) * g;
        }
        const int gn = global_index(pn, nb);
        const R raN = with_a ? one / dom.A[gn] : one;
        if (with_visc) scatter_viscosity<R>(dom, nb, pn, with_a ? raN_grad * raN : raN_grad, b, u);
        if (with_a && dom.A_grad) {
            if (with_visc) raN_grad *= viscosity(dom, nb, pn);
            add_to(dom.A_grad + gn, -raN_grad * raN * raN);
        }
    }
    const int g = global_index(pos, b);
    const R raP = with_a ? one / dom.A[g] : one;
    if (with_visc) scatter_viscosity<R>(dom, b, pos, with_a ? raP_grad * raP : raP_grad, b, u);
    if (with_a && dom.A_grad) {
        if (with_visc) raP_grad *= viscosity(dom, b, pos);
        add_to(dom.A_grad + g, -raP_grad * raP * raP);
    }
}

// VJP of nonortho_matrix_face (PISO_build_matrix_GRAD :3948-4019, PISO_build_pressure_matrix_GRAD :5038-5104):
// accumulates into ai_grad the gradient of the face's cross coefficients from the row gradient cg (slot order,
// cg[0] the diagonal's).
template <typename R, int D>
__device__ __forceinline__ void nonortho_matrix_face_grad(const DomainK<R> &dom, int bi, const I4 &pos, int face,
                                                          const R *ai, int flags, R sign, int corner_type,
                                                          bool grad_always, const R *cg, R *ai_grad) {
    const int dim = face >> 1, fs = face_sign(face);
    const bool nbrs = flags & PICT_NON_ORTHO_DIRECT_MATRIX, ctr = flags & PICT_NON_ORTHO_CENTER_MATRIX;
#pragma unroll
    for (int i = 1; i < D; ++i) {
        const int ta = (dim + i) % D;
        if (ai[face * (D - 1) + i - 1] == R(0)) continue;
        R ag = 0;
#pragma unroll
        for (int tu = 0; tu < 2; ++tu) {
            const int tf = (ta << 1) | tu, tfs = face_sign(tf);
            const Corner<R> cv = corner_value<R, D>(dom, bi, pos, face, tf, false, false, corner_type);
            if (cv.num < 1) {
                if (grad_always || cv.btype == PICT_BC_NEUMANN) {
                    R vg = 0;
                    if (ctr) vg += 3 * cg[0];
                    if (nbrs) {
                        vg += 3 * cg[face + 1];
                        vg -= cg[(tf ^ 1) + 1];
                    }
                    ag += (R)(fs * tfs) * R(0.25) * (sign * vg);
                }
            } else {
                const R norm = (R)(1.0 / (double)cv.num);
                R vg = 0;
                if (ctr) vg += cg[0];
                if (nbrs) {
                    vg += cg[face + 1];
                    vg += cg[tf + 1];
                }
                ag += (R)(fs * tfs) * norm * (sign * vg);
            }
        }
        ai_grad[face * (D - 1) + i - 1] += ag;
    }
}

// scatterNonOrthoLaplaceRHS_v2_GRAD :3210-3380 (cell metrics): VJP of nonortho_rhs for the output gradient s_grad
// into the result gradient of `type`, FixedBoundary velocity gradients, viscosities and A.
template <typename R, int D, bool FACE = false>
__device__ __forceinline__ void nonortho_rhs_grad(const DomainK<R> &dom, int bi, const I4 &pos, int flags, int type,
                                                  bool with_visc, bool with_a, R s_grad, UniformAcc<R> &u) {
    if (!(flags & kNonOrthoRhs)) return;
    const pict_block_desc *b = dom.blocks + bi;
    const bool direct = flags & PICT_NON_ORTHO_DIRECT_RHS;
    R aP[D * (D + 1) / 2];
    lap_full<R, D>(b, pos, aP);
    const R one = 1;
    const int g = global_index(pos, b);
    const R raP = with_a ? one / dom.A[g] : one;
    const R viscP = with_visc ? viscosity(dom, b, pos) : one;
    const R rvP = raP * viscP;
    const bool diffable = with_visc || with_a;
    R raP_grad = 0;
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
                const R tgg = -(R)fs * bA[ta] * s_grad * df;
                if (with_visc) {
                    const R tdg = fixed_data<R>(b, hi, face, type) - fixed_data<R>(b, lo, face, type);
                    scatter_viscosity<R>(dom, b, pos, tgg * tdg, b, u);
                }
                const R bv = with_visc ? viscosity(dom, b, pos) : one;
                scatter_fixed_data<R>(b, hi, face, type, tgg * bv);
                scatter_fixed_data<R>(b, lo, face, type, -tgg * bv);
            }
            continue;
        }
        const NCell n = resolve_neighbor<D>(dom.blocks, bi, pos, face);
        R aN[D];
        lap_single_axis_neighbor<R, D>(dom.blocks, n, face, aN);
        const pict_block_desc *nb = dom.blocks + n.bi;
        const int gn = global_index(n.pos, nb);
        const R raN = with_a ? one / dom.A[gn] : one;
        const R viscN = with_visc ? viscosity(dom, nb, n.pos) : one;
        const R rvN = raN * viscN;
        R raN_grad = 0;
#pragma unroll
        for (int i = 1; i < D; ++i) {
            const int ta = (axis + i) % D;
            // useFaceTransform (:3294-3312, :3364-3368): the face's coefficient for both cells
            const R fc = FACE ? lap_single_face<R, D>(b, pos, face, ta) : R(0);
            const R cP = FACE ? fc : aP[lap_ci(axis, ta)], cN = FACE ? fc : aN[ta];
            const R fa = (cP * rvP + cN * rvN) * R(0.5);
            const R tgg = -s_grad * (R)fs * fa;
            R tdg = 0;
#pragma unroll
            for (int tu = 0; tu < 2; ++tu) {
                const int tf = (ta << 1) | tu, tfs = face_sign(tf);
                const Corner<R> cv = corner_value<R, D>(dom, bi, pos, face, tf, false, direct, type);
                if (cv.num == 0 && (cv.btype == PICT_BC_NEUMANN || type == GD_PRESSURE_RESULT)) {
                    const int to = tf ^ 1, tofs = -tfs;
                    if (direct) {
                        if (diffable) tdg += (R)tfs * data_neighbor<R, D>(dom, bi, pos, face, type) * R(0.75);
                        scatter_data_neighbor<R, D>(dom, bi, pos, face, type, tgg * (R)tfs * R(0.75));
                        if (diffable) tdg += (R)tofs * data_neighbor<R, D>(dom, bi, pos, to, type) * R(0.25);
                        scatter_data_neighbor<R, D>(dom, bi, pos, to, type, tgg * (R)tofs * R(0.25));
                    }
                    if (flags & PICT_NON_ORTHO_DIAGONAL_RHS) {
                        if (diffable) tdg += (R)tofs * data_neighbor_diag<R, D>(dom, bi, pos, face, to, type) * R(0.25);
                        scatter_data_neighbor_diag<R, D>(dom, bi, pos, face, to, type, tgg * (R)tofs * R(0.25));
                    }
                } else {
                    if (diffable) tdg += (R)tfs * cv.data;
                    scatter_corner_value<R, D>(dom, bi, pos, face, tf, direct, type, tgg * (R)tfs, cv.num);
                }
            }
            if (diffable) {
                const R fag = -s_grad * (R)fs * tdg;
                raP_grad += cP * R(0.5) * fag;
                raN_grad += cN * R(0.5) * fag;
            }
        }
        if (with_visc) scatter_viscosity<R>(dom, nb, n.pos, raN_grad * raN, b, u);
        if (with_a && dom.A_grad) add_to(dom.A_grad + gn, raN_grad * viscN * (-raN * raN));
    }
    if (with_visc) scatter_viscosity<R>(dom, b, pos, raP_grad * raP, b, u);
    if (with_a && dom.A_grad) add_to(dom.A_grad + g, raP_grad * viscP * (-raP * raP));
}

// ---------------------------------------------------------------------------
// PISO_build_matrix_GRAD :3869-4075: C_grad, A_grad -> block velocity grad,
// viscosity grads (NO: with the non-orthogonal matrix terms of `flags`).
// ---------------------------------------------------------------------------
// SC: the matrix of passive-scalar channel ch (forPassiveScalar): viscosity gradients go to the channel's scalar
// viscosity (u.visc_scalar), the cross coefficients' gradient times their viscosity-free value (:4004-4011).
template <typename R, int D, bool NO, bool SC = false>
__device__ __forceinline__ void adv_matrix_grad_cell(const DomainK<R> &dom, const pict_block_desc *b, int bi, int f,
                                                     int flags, UniformAcc<R> &u, int ch = 0) {
    const I4 pos = unflat(f, b);
    const int g = b->global_offset + f;
    const R rdet = R(1) / cell_det<R, D>(b, pos);
    int at[7];
    slot_positions<R, D>(dom, b, bi, pos, g, at);
    R cg[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) cg[k] = (at[k] >= 0 && dom.C_grad) ? dom.C_grad[at[k]] * rdet : R(0);
    const R diag_grad = (dom.A_grad ? dom.A_grad[g] * rdet : R(0)) + cg[0];
    R alpha_p[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        I4 t = pos;
        t.a[3] = d;
        alpha_p[d] = laplace_orth<R, D>(b, t);
    }
    R fg[2 * D];
    R visc_grad = 0;
    constexpr int NA = NO && D > 1 ? 2 * D * (D - 1) : 1;
    R ai[NA], ai_grad[NA];
    const bool no = NO && D > 1 && (flags & kNonOrthoMatrix);
    if constexpr (NO && D > 1) {
        if (no) interp_nonortho<R, D>(dom, bi, pos, false, false, ai);
#pragma unroll
        for (int k = 0; k < NA; ++k) ai_grad[k] = 0;
    }
#pragma unroll
    for (int face = 0; face < 2 * D; ++face) {
        const int dim = face >> 1;
        const R fs = (face & 1) ? R(1) : R(-1);
        if (!(at_bound(pos, face, b) && prescribed(face, b))) {
            int nb;
            I4 t = neighbor_cell<D>(dom.blocks, b, pos, face, bi, nb);
            const pict_block_desc *nbp = dom.blocks + nb;
            const R alpha = laplace_orth<R, D>(nbp, t);
            fg[face] = fs * R(0.5) * (diag_grad + cg[face + 1]);
            const R vc_grad = diag_grad - cg[face + 1];
            visc_grad += R(0.5) * alpha_p[dim] * vc_grad;
            t.a[3] = 0;
            if constexpr (SC) u.visc_scalar += (double)(R(0.5) * alpha * vc_grad);
            else scatter_viscosity<R>(dom, nbp, t, R(0.5) * alpha * vc_grad, b, u);
            if constexpr (NO && D > 1) {
                if (no) {
                    R cgd[7];
#pragma unroll
                    for (int k = 0; k < 7; ++k) cgd[k] = cg[k];
                    cgd[0] = diag_grad;
                    I4 cp = pos;
                    if constexpr (SC) cp.a[3] = ch;
                    nonortho_matrix_face_grad<R, D>(dom, bi, cp, face, ai, flags, R(-1), SC ? GD_SCALAR : GD_VELOCITY,
                                                    false, cgd, ai_grad);
                }
            }
        } else {
            fg[face] = 0;
            const int bc = SC ? fixed_scalar_bc(&b->bounds[face], ch) : b->bounds[face].velocity_bc;
            const R slip = bc == PICT_BC_DIRICHLET ? R(0) : R(1);
            visc_grad += (R(1) - slip) * R(2) * alpha_p[dim] * diag_grad;
        }
    }
    if constexpr (NO && D > 1) {
        if (no) {
            if constexpr (SC) {
#pragma unroll
                for (int k = 0; k < NA; ++k) visc_grad += ai_grad[k] * ai[k];
            } else {
                scatter_interp_nonortho<R, D>(dom, bi, pos, ai_grad, true, false, u);
            }
        }
    }
    if (b->velocity_grad) scatter_fluxes<R, D>(dom, b, pos, fg, nullptr);
    if constexpr (SC) u.visc_scalar += (double)visc_grad;
    else scatter_viscosity<R>(dom, b, pos, visc_grad, b, u);
}

template <typename R, int D, bool NO>
__global__ __launch_bounds__(kCellsPerTile) void k_adv_matrix_scalar_grad(DomainK<R> dom, R dt, int flags, int ch) {
    const pict_block_desc *b;
    int bi, f;
    (void)dt;
    UniformAcc<R> u;
    if (tile_cell(dom, b, bi, f)) adv_matrix_grad_cell<R, D, NO, true>(dom, b, bi, f, flags, u, ch);
    flush_scalar_viscosity<R>(dom, u.visc_scalar, ch);
}

// kPISO_build_scalar_advection_RHS_GRAD :4193-4278: scalar_rhs_grad -> block scalar grad (=), fixed-boundary scalar
// and velocity grads (+=), scalar viscosity grad (+=) (NO: and the VJP of the explicit non-orthogonal terms of
// `flags`, into scalar_result_grad).  Every thread runs the channel loop (the viscosity sums are workgroup sums).
template <typename R, int D, bool NO>
__global__ __launch_bounds__(kCellsPerTile) void k_scalar_rhs_grad(DomainK<R> dom, R dt, int flags) {
    const pict_block_desc *b;
    int bi, f;
    const bool live = tile_cell(dom, b, bi, f);
    UniformAcc<R> u;  // unused: the explicit terms of a scalar carry no viscosity of their own
    const I4 pos = live ? unflat(f, b) : I4{{0, 0, 0, 0}};
    const R det = live ? cell_det<R, D>(b, pos) : R(1);
    for (int c = 0; c < dom.channels; ++c) {
        double vacc = 0.0;
        if (live) {
            I4 t = pos;
            t.a[3] = c;
            const int g = b->global_offset + f;
            const R visc = scalar_viscosity(dom, c);
            R G = dom.srhs_grad[(size_t)c * dom.N + g] / det;
            if constexpr (NO && D > 1) {
                nonortho_rhs_grad<R, D>(dom, bi, t, flags, GD_SCALAR_RESULT, false, false, -G * visc, u);
                vacc += (double)(-G * nonortho_rhs<R, D>(dom, bi, t, flags, GD_SCALAR_RESULT, false, false));
            }
#pragma unroll
            for (int face = 0; face < 2 * D; ++face) {
                if (!at_bound(pos, face, b)) continue;
                const pict_boundary_desc *fb = &b->bounds[face];
                if (fb->type != PICT_BOUNDARY_FIXED) continue;
                const int ax = face >> 1;
                const R normal = (face & 1) ? R(1) : R(-1);
                I4 bp = pos;
                bp.a[ax] = 0;
                const R sv = fixed_scalar<R>(fb, bp, c);
                const bool dirichlet = fixed_scalar_bc(fb, c) == PICT_BC_DIRICHLET;
                bp.a[3] = ax;
                const R alpha = laplace_orth_fixed<R, D>(fb, bp);
                const R flux = fixed_contravariant<R, D>(fb, bp) * normal;
                R s_grad = -flux * G;
                if (dirichlet) {
                    s_grad += visc * R(2) * alpha * G;
                    vacc += (double)(sv * R(2) * alpha * G);
                } else {
                    s_grad += visc * G;
                    vacc += (double)(sv * G);
                }
                scatter_fixed_contravariant<R, D>(fb, bp, -sv * G * normal);
                scatter_fixed_data<R>(b, t, face, GD_SCALAR, s_grad);
            }
            if (b->scalar_grad) reinterpret_cast<R *>(b->scalar_grad)[flat4(t, b->stride)] = G * det / dt;
        }
        flush_scalar_viscosity<R>(dom, vacc, c);
    }
}

template <typename R, int D, bool NO>
__global__ __launch_bounds__(kCellsPerTile) void k_adv_matrix_grad(DomainK<R> dom, R dt, int flags) {
    const pict_block_desc *b;
    int bi, f;
    (void)dt;
    UniformAcc<R> u;
    if (tile_cell(dom, b, bi, f)) adv_matrix_grad_cell<R, D, NO>(dom, b, bi, f, flags, u);
    flush_uniform<R, D>(dom, b, u, false);
}

// kPISO_build_advection_RHS_GRAD :4388-4476: velocity_rhs_grad -> velocity grad (=),
// velocity source grad, fixed boundary velocity and viscosity grads (NO: and the VJP of the
// explicit non-orthogonal terms of `flags`, into velocity_result_grad).
template <typename R, int D, bool NO>
__global__ __launch_bounds__(kCellsPerTile) void k_adv_rhs_grad(DomainK<R> dom, R dt, int flags) {
    const pict_block_desc *b;
    int bi, f;
    UniformAcc<R> u;
    if (tile_cell(dom, b, bi, f)) {
        const I4 pos = unflat(f, b);
        const int g = b->global_offset + f;
        const R det = cell_det<R, D>(b, pos);
        for (int d = 0; d < D; ++d) {
            I4 t = pos;
            t.a[3] = d;
            R G = dom.vrhs_grad[(size_t)d * dom.N + g];
            scatter_velocity_source<R>(b, t, G, u);
            G /= det;
            if constexpr (NO && D > 1) nonortho_rhs_grad<R, D>(dom, bi, t, flags, GD_VELOCITY_RESULT, true, false, -G, u);
            fixed_boundary_source_grad<R, D>(dom, b, pos, d, G, u);
            G *= det;
            if (b->velocity_grad) reinterpret_cast<R *>(b->velocity_grad)[flat4(t, b->stride)] = G / dt;
        }
    }
    flush_uniform<R, D>(dom, b, u, true);
}

// PISO_build_pressure_matrix_GRAD :4967-5115: P_grad -> A_grad (+=) (NO: with the non-orthogonal
// matrix terms of `flags`).
template <typename R, int D, bool NO>
__global__ __launch_bounds__(kCellsPerTile) void k_p_matrix_grad(DomainK<R> dom, int flags) {
    const pict_block_desc *b;
    int bi, f;
    if (!tile_cell(dom, b, bi, f)) return;
    const I4 pos = unflat(f, b);
    const int g = b->global_offset + f;
    int at[7];
    slot_positions<R, D>(dom, b, bi, pos, g, at);
    const R pg0 = dom.P_grad[at[0]];
    R raP_grad = 0;
    constexpr int NA = NO && D > 1 ? 2 * D * (D - 1) : 1;
    R ai[NA], ai_grad[NA], pcg[7];
    const bool no = NO && D > 1 && (flags & kNonOrthoMatrix);
    if constexpr (NO && D > 1) {
        if (no) {
            interp_nonortho<R, D>(dom, bi, pos, false, false, ai);
#pragma unroll
            for (int k = 0; k < 7; ++k) pcg[k] = at[k] >= 0 ? dom.P_grad[at[k]] : R(0);
        }
#pragma unroll
        for (int k = 0; k < NA; ++k) ai_grad[k] = 0;
    }
#pragma unroll
    for (int face = 0; face < 2 * D; ++face) {
        if (at_bound(pos, face, b) && prescribed(face, b)) continue;
        if constexpr (NO && D > 1) {
            if (no) nonortho_matrix_face_grad<R, D>(dom, bi, pos, face, ai, flags, R(1), GD_IS_FIXED, true, pcg, ai_grad);
        }
        const int dim = face >> 1;
        int nb;
        I4 t = neighbor_cell<D>(dom.blocks, b, pos, face, bi, nb);
        const pict_block_desc *nbp = dom.blocks + nb;
        const R alphaN = laplace_orth<R, D>(nbp, t);
        I4 tp = pos;
        tp.a[3] = dim;
        const R alphaP = laplace_orth<R, D>(b, tp);
        t.a[3] = 0;
        const int gn = global_index(t, nbp);
        const R raN = R(1) / dom.A[gn];
        const R coef_grad = dom.P_grad[at[face + 1]] - pg0;
        raP_grad += R(0.5) * alphaP * coef_grad;
        add_to(dom.A_grad + gn, -(R(0.5) * alphaN * coef_grad) * raN * raN);
    }
    const R raP = R(1) / dom.A[g];
    add_to(dom.A_grad + g, -raP_grad * raP * raP);
    if constexpr (NO && D > 1) {
        if (no) {
            UniformAcc<R> u;  // no viscosity targets here
            scatter_interp_nonortho<R, D>(dom, bi, pos, ai_grad, false, true, u);
        }
    }
}

// k_pressureRHSaddNonOrthoComponents_GRAD :5479-5494: pressure_rhs_div_grad -> pressure_result_grad (+=), A_grad (+=)
// (FACE: with useFaceTransform's face coefficients, which this VJP does use, :5491)
template <typename R, int D, bool FACE = false>
__global__ __launch_bounds__(kCellsPerTile) void k_p_rhs_nonortho_grad(DomainK<R> dom, int flags) {
    const pict_block_desc *b;
    int bi, f;
    if (!tile_cell(dom, b, bi, f)) return;
    UniformAcc<R> u;  // no viscosity targets here
    if constexpr (D > 1)
        nonortho_rhs_grad<R, D, FACE>(dom, bi, unflat(f, b), flags, GD_PRESSURE_RESULT, false, true,
                                dom.prhs_div_grad[b->global_offset + f], u);
}

// k_computePressureRHSdivergenceFromFlux_GRAD :5423-5452: pressure_rhs_div_grad ->
// pressure_rhs_grad (+=), fixed boundary velocity grads.
template <typename R, int D>
__global__ __launch_bounds__(kCellsPerTile) void k_div_grad(DomainK<R> dom, R dt, int tsn) {
    const pict_block_desc *b;
    int bi, f;
    if (!tile_cell(dom, b, bi, f)) return;
    const I4 pos = unflat(f, b);
    R dg = dom.prhs_div_grad[b->global_offset + f];
    if (tsn) dg /= dt;
    R fg[2 * D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        fg[2 * d + 1] = dg;
        fg[2 * d] = -dg;
    }
    scatter_fluxes<R, D>(dom, b, pos, fg, dom.prhs_grad);
}

// PISO_build_pressure_rhs_GRAD :5243-5369: pressure_rhs_grad -> velocity grad (=),
// velocity_result grad (+=), C_grad (= over the row), A_grad (+=), source and
// boundary grads.
template <typename R, int D>
__device__ __forceinline__ void p_rhs_grad_cell(const DomainK<R> &dom, const pict_block_desc *b, int f, R dt,
                                                UniformAcc<R> &u) {
    const I4 pos = unflat(f, b);
    const int g = b->global_offset + f;
    const R rDiag = R(1) / dom.A[g];
    const R det = cell_det<R, D>(b, pos);
    const int cs = dom.csr_row[g], ce = dom.csr_row[g + 1];
    const int rs = ce - cs < 7 ? ce - cs : 7;
    int col[7];
    R val[7], cvg[7];
    for (int k = 0; k < rs; ++k) {
        col[k] = dom.csr_index[cs + k];
        val[k] = dom.C[cs + k];
        cvg[k] = 0;
    }
    R rDiag_grad = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        I4 t = pos;
        t.a[3] = d;
        const size_t o = (size_t)dom.N * d;
        const R pGrad = dom.prhs_grad[g + o];
        rDiag_grad += dom.prhs[g + o] / rDiag * pGrad;
        if (b->velocity_grad) reinterpret_cast<R *>(b->velocity_grad)[flat4(t, b->stride)] = rDiag / dt * pGrad;
        const R w = rDiag * pGrad;
        for (int k = 0; k < rs; ++k) {
            if (col[k] == g) continue;
            if (dom.vres_grad) add_to(dom.vres_grad + col[k] + o, -val[k] * w);
            cvg[k] -= dom.vres[col[k] + o] * w;
        }
        R S_grad = w;
        scatter_velocity_source<R>(b, t, S_grad, u);
        S_grad /= det;
        fixed_boundary_source_grad<R, D>(dom, b, pos, d, S_grad, u);
    }
    if (dom.C_grad)
        for (int k = 0; k < rs; ++k) dom.C_grad[cs + k] = cvg[k];
    if (dom.A_grad) dom.A_grad[g] += rDiag_grad * (-rDiag * rDiag);
}

template <typename R, int D>
__global__ __launch_bounds__(kCellsPerTile) void k_p_rhs_grad(DomainK<R> dom, R dt) {
    const pict_block_desc *b;
    int bi, f;
    UniformAcc<R> u;
    if (tile_cell(dom, b, bi, f)) p_rhs_grad_cell<R, D>(dom, b, f, dt, u);
    flush_uniform<R, D>(dom, b, u, true);
}

// scatterPressureGradToWithBounds :3411-3477 (transpose of pressure_at)
template <typename R, int D>
__device__ __forceinline__ void scatter_pressure_to(const DomainK<R> &dom, const pict_block_desc *b, const I4 &pos,
                                                    R g) {
    I4 t = pos;
    t.a[3] = 0;
    R *own = reinterpret_cast<R *>(b->pressure_grad);
#pragma unroll
    for (int dim = 0; dim < D; ++dim) {
        if (pos.a[dim] < 0 || pos.a[dim] >= b->size[dim]) {
            const bool lower = pos.a[dim] < 0;
            const pict_boundary_desc *bd = &b->bounds[2 * dim + (lower ? 0 : 1)];
            if (bd->type == PICT_BOUNDARY_FIXED) {
                t.a[dim] = lower ? 0 : b->size[dim] - 1;
                if (own) add_to(own + flat4(t, b->stride), g);
            } else if (bd->type == PICT_BOUNDARY_CONNECTED) {
                const pict_block_desc *ob = dom.blocks + bd->connected_block;
                const int border = lower ? -(pos.a[dim] + 1) : pos.a[dim] - b->size[dim];
                I4 o = connected_pos<D>(dom.blocks, t, dim, bd, border);
                if (ob->pressure_grad) add_to(reinterpret_cast<R *>(ob->pressure_grad) + flat4(o, ob->stride), g);
            } else if (bd->type == PICT_BOUNDARY_PERIODIC) {
                t.a[dim] += lower ? b->size[dim] : -b->size[dim];
                if (own) add_to(own + flat4(t, b->stride), g);
            }
            return;
        }
    }
    if (own) add_to(own + flat4(t, b->stride), g);
}

// PISO_update_velocity_GRAD :6250-6293 (+ scatterPressureGradientGrad :838-870):
// velocity_result_grad -> pressure_rhs_grad (=), block pressure grads (+=), A_grad (=).
template <typename R, int D>
__global__ __launch_bounds__(kCellsPerTile) void k_correct_velocity_grad(DomainK<R> dom, R dt, int tsn) {
    const pict_block_desc *b;
    int bi, f;
    if (!tile_cell(dom, b, bi, f)) return;
    const I4 pos = unflat(f, b);
    const int g = b->global_offset + f;
    const R rDiag = R(1) / dom.A[g];
    R pg[D], pgg[D];
    pressure_gradient<R, D>(dom, b, pos, pg);
    R rDiag_grad = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const size_t o = (size_t)dom.N * d;
        R vug = dom.vres_grad[g + o];
        dom.prhs_grad[g + o] = vug;
        if (tsn) vug *= dt;
        pgg[d] = -rDiag * vug;
        rDiag_grad -= pg[d] * vug;
    }
    {
        R v[D];
        if (b->has_transform) {
            const R *T = cell_T<R, D>(b, pos);
#pragma unroll
            for (int i = 0; i < D; ++i) {
                R s = 0;
#pragma unroll
                for (int j = 0; j < D; ++j) s += Tminv<R, D>(T, i, j) * pgg[j];
                v[i] = s;
            }
        } else {
#pragma unroll
            for (int i = 0; i < D; ++i) v[i] = pgg[i];
        }
#pragma unroll
        for (int dim = 0; dim < D; ++dim) {
            I4 pP = pos, pN = pos;
            pP.a[3] = pN.a[3] = 0;
            R fac = R(0.5);
            if (!(pos.a[dim] == 0 && prescribed(2 * dim, b))) pN.a[dim] = pos.a[dim] - 1;
            else fac = R(1);
            if (!(pos.a[dim] == b->size[dim] - 1 && prescribed(2 * dim + 1, b))) pP.a[dim] = pos.a[dim] + 1;
            else fac = R(1);
            scatter_pressure_to<R, D>(dom, b, pN, v[dim] * -fac);
            scatter_pressure_to<R, D>(dom, b, pP, v[dim] * fac);
        }
    }
    dom.A_grad[g] = -rDiag_grad * rDiag * rDiag;
}

}  // namespace pict

using namespace pict;

extern "C" {

pict_status pict_setup_advection_matrix_grad(const pict_atlas *a, double dt, int32_t flags, void *stream) {
    PICT_CHECK_ARG(a && a->desc.C_grad && a->desc.A_grad, "C_grad/A_grad missing");
    if (a->desc.num_dims > 1 && (flags & kNonOrthoMatrix))
        PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_adv_matrix_grad<R, D, true>), a->view<R>(), (R)dt,
                                           (int)flags));
    else
        PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_adv_matrix_grad<R, D, false>), a->view<R>(), (R)dt, 0));
    return check_launch();
}

pict_status pict_setup_advection_matrix_scalar_grad(const pict_atlas *a, double dt, int32_t flags, int32_t channel,
                                                    void *stream) {
    PICT_CHECK_ARG(a && a->desc.C_grad && a->desc.A_grad, "C_grad/A_grad missing");
    PICT_CHECK_ARG(a->desc.scalar_channels > 0, "domain has no passive scalar");
    PICT_CHECK_ARG(channel >= 0 && channel < a->desc.scalar_channels, "passive scalar channel %d out of range",
                   channel);
    if (a->desc.num_dims > 1 && (flags & kNonOrthoMatrix))
        PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_adv_matrix_scalar_grad<R, D, true>), a->view<R>(), (R)dt,
                                           (int)flags, (int)channel));
    else
        PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_adv_matrix_scalar_grad<R, D, false>), a->view<R>(), (R)dt,
                                           0, (int)channel));
    return check_launch();
}

pict_status pict_setup_advection_scalar_grad(const pict_atlas *a, double dt, int32_t flags, void *stream) {
    PICT_CHECK_ARG(a && a->desc.scalar_channels > 0, "domain has no passive scalar");
    PICT_CHECK_ARG(a->desc.scalar_rhs_grad, "scalar_rhs_grad missing");
    for (int i = 0; i < a->desc.num_blocks; ++i)
        PICT_CHECK_ARG(a->host_blocks[i].scalar, "block %d: passive scalar missing", i);
    if (a->desc.num_dims > 1 && (flags & kNonOrthoRhs)) {
        PICT_CHECK_ARG(a->desc.scalar_result, "scalar_result missing (non-orthogonal RHS terms)");
        PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_scalar_rhs_grad<R, D, true>), a->view<R>(), (R)dt,
                                           (int)flags));
    } else {
        PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_scalar_rhs_grad<R, D, false>), a->view<R>(), (R)dt, 0));
    }
    return check_launch();
}

pict_status pict_setup_advection_velocity_grad(const pict_atlas *a, double dt, int32_t flags, int32_t apply_pg,
                                               void *stream) {
    (void)apply_pg;  // the reference's VJP omits the pressure-gradient term (:4388-4476)
    PICT_CHECK_ARG(a && a->desc.velocity_rhs_grad, "velocity_rhs_grad missing");
    if (a->desc.num_dims > 1 && (flags & kNonOrthoRhs)) {
        PICT_CHECK_ARG(a->desc.velocity_result && a->desc.velocity_result_grad,
                       "velocity_result / its gradient missing (non-orthogonal RHS terms)");
        PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_adv_rhs_grad<R, D, true>), a->view<R>(), (R)dt, (int)flags));
    } else {
        PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_adv_rhs_grad<R, D, false>), a->view<R>(), (R)dt, 0));
    }
    return check_launch();
}

pict_status pict_setup_pressure_matrix_grad(const pict_atlas *a, double dt, int32_t flags, int32_t use_face_transform,
                                            void *stream) {
    (void)dt;
    PICT_CHECK_ARG(a && a->desc.P_grad && a->desc.A_grad && a->desc.A, "P_grad/A_grad missing");
    // PISO_build_pressure_matrix_GRAD (:4966-5115) takes useFaceTransform but passes false to its interpolation and
    // scatter of the non-orthogonal coefficients (:5043-5044, :5104-5105): the cell-metric VJP either way
    (void)use_face_transform;
    if (a->desc.num_dims > 1 && (flags & kNonOrthoMatrix))
        PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_p_matrix_grad<R, D, true>), a->view<R>(), (int)flags));
    else
        PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_p_matrix_grad<R, D, false>), a->view<R>(), 0));
    return check_launch();
}

pict_status pict_setup_pressure_rhs_div_grad(const pict_atlas *a, double dt, int32_t flags, int32_t use_face_transform,
                                             int32_t tsn, void *stream) {
    PICT_CHECK_ARG(a && a->desc.pressure_rhs_div_grad && a->desc.pressure_rhs_grad, "pressure rhs grads missing");
    // k_computePressureRHSdivergenceFromFlux_GRAD (:5422-5452) scatters through the cell-metric fluxes whatever
    // useFaceTransform says; the non-orthogonal terms' VJP uses the face coefficients (:5491)
    if (a->desc.num_dims > 1 && (flags & kNonOrthoRhs)) {
        // k_pressureRHSaddNonOrthoComponents_GRAD first (_SetupPressureRHS_GRAD :5792-5803)
        PICT_CHECK_ARG(a->desc.pressure_result && a->desc.pressure_result_grad && a->desc.A && a->desc.A_grad,
                       "pressure_result / its gradient or A_grad missing (non-orthogonal RHS terms)");
        if (use_face_transform) {
            for (int i = 0; i < a->desc.num_blocks; ++i)
                PICT_CHECK_ARG(!a->host_blocks[i].has_transform ||
                                   (a->host_blocks[i].has_face_transform && a->host_blocks[i].face_transform),
                               "useFaceTransform with non-orthogonal terms: block %d has no face transforms", i);
            PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_p_rhs_nonortho_grad<R, D, true>), a->view<R>(),
                                               (int)flags));
        } else {
            PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_p_rhs_nonortho_grad<R, D>), a->view<R>(), (int)flags));
        }
        pict_status st = check_launch();
        if (st) return st;
    }
    PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_div_grad<R, D>), a->view<R>(), (R)dt, (int)tsn));
    return check_launch();
}

pict_status pict_setup_pressure_rhs_grad(const pict_atlas *a, double dt, int32_t flags, int32_t use_face_transform,
                                         int32_t tsn, void *stream) {
    PICT_CHECK_ARG(a && a->desc.pressure_rhs_grad && a->desc.pressure_rhs && a->desc.C_value &&
                       a->desc.velocity_result && a->desc.A,
                   "pressure rhs grad inputs missing");
    pict_status st = pict_setup_pressure_rhs_div_grad(a, dt, flags, use_face_transform, tsn, stream);
    if (st) return st;
    PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_p_rhs_grad<R, D>), a->view<R>(), (R)dt));
    return check_launch();
}

pict_status pict_setup_pressure_correction_grad(const pict_atlas *a, double dt, int32_t flags,
                                                int32_t use_face_transform, int32_t tsn, void *stream) {
    pict_status st = pict_setup_pressure_rhs_grad(a, dt, flags, use_face_transform, tsn, stream);
    if (st) return st;
    return pict_setup_pressure_matrix_grad(a, dt, flags, use_face_transform, stream);
}

pict_status pict_correct_velocity_grad(const pict_atlas *a, double dt, int32_t tsn, void *stream) {
    PICT_CHECK_ARG(a && a->desc.velocity_result_grad && a->desc.pressure_rhs_grad && a->desc.A_grad && a->desc.A,
                   "correct velocity grads missing");
    PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_correct_velocity_grad<R, D>), a->view<R>(), (R)dt, (int)tsn));
    return check_launch();
}

}  // extern "C"
