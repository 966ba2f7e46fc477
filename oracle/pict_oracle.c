/*
 * pict_oracle.c -- CPU restatement of the PICT PISO time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for libpict_hip: only
 * tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py load
 * it.  The product path never calls it.
 *
 * It reads the same pict_domain_desc as the HIP library (include/pict_hip.h),
 * with host pointers.  Built twice:
 *   liboracle_f64.so  (ORACLE_REAL=double): the FP64 oracle
 *   liboracle_f32.so  (ORACLE_REAL=float) : the FP32 CPU baseline ("port")
 * Every function names the reference lines it restates
 * (tum-pbs/PICT, extensions/PISO_multiblock_cuda_kernel.cu unless stated).
 *
 * Pinned (the reference ships no golden vectors) by checks that do not restate
 * these formulas: closed forms and brute force for the CSR pattern; scipy for
 * the solvers; finite differences for every VJP; for the pressure projection
 * the divergence of linear fields, the continuum limit of the pressure matrix,
 * the exact divergence left by one projection of a Fourier mode, the
 * reference's step invariants and the Re 100 cavity against Ghia et al.
 * (tests/test_oracle*.py; tests/test_oracle_mutations.py plants mistakes in
 * copies of this file and shows the pins catch them).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/pict_hip.h"

#ifndef ORACLE_REAL
#define ORACLE_REAL double
#endif
typedef ORACLE_REAL real;

#define ORACLE_DTYPE ((sizeof(real) == 8) ? PICT_DTYPE_F64 : PICT_DTYPE_F32)

typedef struct { int a[4]; } I4;

static char g_err[512];
const char *oracle_last_error(void) { return g_err; }
int oracle_real_bytes(void) { return (int)sizeof(real); }

void oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------ */
/* index helpers: flattenIndex / unflattenIndex, grid_definitions.h    */
/* ------------------------------------------------------------------ */

static inline int flat4(const I4 p, const int32_t *stride) {
    return p.a[0] + stride[1] * p.a[1] + stride[2] * p.a[2] + stride[3] * p.a[3];
}
static inline I4 unflat(int idx, const pict_block_desc *b) {
    I4 p;
    p.a[0] = idx % b->size[0];
    p.a[1] = (idx / b->stride[1]) % b->size[1];
    p.a[2] = (idx / b->stride[2]) % b->size[2];
    p.a[3] = 0;
    return p;
}
static inline int posmod(int a, int m) { int r = a % m; return r < 0 ? r + m : r; }

/* isAtBound :183-193 */
static inline int at_bound(const I4 p, int face, const pict_block_desc *b) {
    int ax = face >> 1;
    return (face & 1) ? p.a[ax] == b->size[ax] - 1 : p.a[ax] == 0;
}
/* isEmptyBound :159-166 (only FIXED is a live prescribed type) */
static inline int prescribed(int face, const pict_block_desc *b) {
    return b->bounds[face].type == PICT_BOUNDARY_FIXED;
}
static inline int global_index(const I4 p, const pict_block_desc *b) {
    I4 q = p; q.a[3] = 0;
    return flat4(q, b->stride) + b->global_offset;
}

typedef struct {
    const pict_domain_desc *d;
    int D, N, TS; /* dims, cells, transform stride 2D^2+1 */
} Ctx;

static inline const pict_block_desc *blk(const Ctx *c, int i) { return &c->d->blocks[i]; }

/* Transform pointer of block cell (w ignored). */
static inline const real *cell_T(const Ctx *c, const pict_block_desc *b, I4 p) {
    p.a[3] = 0;
    return (const real *)b->transform + (size_t)flat4(p, b->stride) * c->TS;
}
static inline const real *bound_T(const Ctx *c, const pict_boundary_desc *fb, I4 p) {
    p.a[3] = 0;
    return (const real *)fb->transform + (size_t)flat4(p, fb->stride) * c->TS;
}
#define T_MINV(c, T, i, j) ((T)[(c)->D * (c)->D + (i) * (c)->D + (j)])
#define T_DET(c, T) ((T)[2 * (c)->D * (c)->D])

/* computeConnectedPos :325-347 */
static I4 connected_pos(const Ctx *c, I4 pos, int bdim, const pict_boundary_desc *cb, int border) {
    const pict_block_desc *ob = blk(c, cb->connected_block);
    I4 o = pos;
    int ca = cb->axes[0] >> 1;
    o.a[ca] = (cb->axes[0] & 1) ? ob->size[ca] - 1 - border : border;
    if (c->D > 1) {
        int ax = (bdim + 1) % c->D;
        ca = cb->axes[1] >> 1;
        o.a[ca] = (cb->axes[1] & 1) ? ob->size[ca] - 1 - pos.a[ax] : pos.a[ax];
        if (c->D > 2) {
            ax = (bdim + 2) % c->D;
            ca = cb->axes[2] >> 1;
            o.a[ca] = (cb->axes[2] & 1) ? ob->size[ca] - 1 - pos.a[ax] : pos.a[ax];
        }
    }
    return o;
}
/* computeConnectedPosWithChannel :351-360 */
static I4 connected_pos_ch(const Ctx *c, I4 pos, int bdim, const pict_boundary_desc *cb) {
    I4 o = connected_pos(c, pos, bdim, cb, 0);
    o.a[3] = cb->axes[posmod(pos.a[3] - bdim, c->D)] >> 1;
    return o;
}

/* Neighbour of a cell across a non-prescribed face, with the channel (w)
   remapped across connections: the neighbour resolution used by
   PISO_build_matrix :3716-3729 and PISO_build_pressure_matrix :4835-4845. */
static I4 neighbor_cell(const Ctx *c, const pict_block_desc *b, I4 pos, int face, int *nb_out, int self_idx) {
    int dim = face >> 1, up = face & 1;
    I4 t = pos; t.a[3] = dim;
    *nb_out = self_idx;
    if (at_bound(pos, face, b) && b->bounds[face].type == PICT_BOUNDARY_CONNECTED) {
        *nb_out = b->bounds[face].connected_block;
        t = connected_pos_ch(c, t, dim, &b->bounds[face]);
    } else if (at_bound(pos, face, b) && b->bounds[face].type == PICT_BOUNDARY_PERIODIC) {
        t.a[dim] = up ? 0 : b->size[dim] - 1;
    } else {
        t.a[dim] = pos.a[dim] + (up ? 1 : -1);
    }
    return t;
}

/* velocity component d of a block cell, block field or global vector :577-593 */
static inline real vel_comp(const Ctx *c, const pict_block_desc *b, I4 p, int d, const real *vg) {
    p.a[3] = d;
    if (!vg) return ((const real *)b->velocity)[flat4(p, b->stride)];
    return vg[(size_t)d * c->N + global_index(p, b)];
}

/* getContravariantComponent :665-678 (p.w = component) */
static real contravariant(const Ctx *c, const pict_block_desc *b, I4 p, const real *vg) {
    int comp = p.a[3];
    if (b->has_transform) {
        const real *T = cell_T(c, b, p);
        real s = 0;
        for (int j = 0; j < c->D; ++j) s += T_MINV(c, T, comp, j) * vel_comp(c, b, p, j, vg);
        return T_DET(c, T) * s;
    }
    return vel_comp(c, b, p, comp, vg);
}

/* FixedBoundary velocity value, getFixedBoundaryData :2419-2437 (VELOCITY) */
static inline real fixed_vel(const Ctx *c, const pict_boundary_desc *fb, I4 p, int d) {
    (void)c;
    if (fb->velocity_static) return ((const real *)fb->velocity)[d];
    p.a[3] = d;
    return ((const real *)fb->velocity)[flat4(p, fb->stride)];
}

/* getContravariantComponentBoundaryFixed :690-710 (p on the boundary grid, w = component) */
static real fixed_contravariant(const Ctx *c, const pict_boundary_desc *fb, I4 p) {
    int comp = p.a[3];
    if (fb->has_transform) {
        const real *T = bound_T(c, fb, p);
        real s = 0;
        /* getVelocityFromBoundaryFixed :611-628: only Dirichlet velocity is read */
        if (fb->velocity_bc == PICT_BC_DIRICHLET)
            for (int j = 0; j < c->D; ++j) s += T_MINV(c, T, comp, j) * fixed_vel(c, fb, p, j);
        return T_DET(c, T) * s;
    }
    return fixed_vel(c, fb, p, comp);
}

/* getLaplaceCoefficientOrthogonal :1209-1224 (p.w = component) */
static real laplace_orth(const Ctx *c, const pict_block_desc *b, I4 p) {
    if (!b->has_transform) return 1;
    const real *T = cell_T(c, b, p);
    real s = 0;
    for (int j = 0; j < c->D; ++j) s += T_MINV(c, T, p.a[3], j) * T_MINV(c, T, p.a[3], j);
    return T_DET(c, T) * s;
}
/* getLaplaceCoefficientOrthogonalBoundary :1454-1466 */
static real laplace_orth_fixed(const Ctx *c, const pict_boundary_desc *fb, I4 p) {
    if (!fb->has_transform) return 1;
    const real *T = bound_T(c, fb, p);
    real s = 0;
    for (int j = 0; j < c->D; ++j) s += T_MINV(c, T, p.a[3], j) * T_MINV(c, T, p.a[3], j);
    return T_DET(c, T) * s;
}

/* getViscosityBlock :1802-1821 (velocity) */
static real viscosity(const Ctx *c, const pict_block_desc *b, I4 p) {
    if (!b->viscosity) return (real)c->d->viscosity;
    if (b->viscosity_static) return ((const real *)b->viscosity)[0];
    p.a[3] = 0;
    return ((const real *)b->viscosity)[flat4(p, b->stride)];
}

/* getBlockVelocitySource :2241-2252 */
static real velocity_source(const pict_block_desc *b, I4 p) {
    if (!b->velocity_source) return 0;
    if (b->velocity_source_static) return ((const real *)b->velocity_source)[p.a[3]];
    return ((const real *)b->velocity_source)[flat4(p, b->stride)];
}

/* getDeterminant :1100-1110 */
static inline real cell_det(const Ctx *c, const pict_block_desc *b, I4 p) {
    return b->has_transform ? T_DET(c, cell_T(c, b, p)) : (real)1;
}

/* computeFluxesNDLoop :1552-1630 -- fluxes NOT multiplied by the face sign */
static void fluxes_nd(const Ctx *c, int bi, I4 pos, real *fl, const real *vg) {
    const pict_block_desc *b = blk(c, bi);
    for (int face = 0; face < 2 * c->D; ++face) {
        int dim = face >> 1, up = face & 1;
        I4 t = pos; t.a[3] = dim;
        real velC = contravariant(c, b, t, vg);
        if (at_bound(pos, face, b)) {
            const pict_boundary_desc *bd = &b->bounds[face];
            switch (bd->type) {
            case PICT_BOUNDARY_FIXED:
                t.a[dim] = 0;
                fl[face] = fixed_contravariant(c, bd, t);
                break;
            case PICT_BOUNDARY_CONNECTED: {
                const pict_block_desc *ob = blk(c, bd->connected_block);
                I4 o = connected_pos(c, t, dim, bd, 0);
                o.a[3] = bd->axes[0] >> 1;
                real velN = contravariant(c, ob, o, vg);
                if ((bd->axes[0] & 1) == up) velN = -velN;
                fl[face] = (velN + velC) * (real)0.5;
                break;
            }
            case PICT_BOUNDARY_PERIODIC:
                t.a[dim] = up ? 0 : b->size[dim] - 1;
                fl[face] = (contravariant(c, b, t, vg) + velC) * (real)0.5;
                break;
            default:
                fl[face] = 0;
            }
        } else {
            t.a[dim] = pos.a[dim] + (up ? 1 : -1);
            fl[face] = (contravariant(c, b, t, vg) + velC) * (real)0.5;
        }
    }
}

static int init_ctx(Ctx *c, const pict_domain_desc *d) {
    if (!d || d->dtype != ORACLE_DTYPE) {
        snprintf(g_err, sizeof g_err, "oracle: descriptor dtype %d does not match oracle real (%d bytes)",
                 d ? d->dtype : -1, (int)sizeof(real));
        return PICT_ERR_INVALID;
    }
    c->d = d; c->D = d->num_dims; c->N = d->num_cells; c->TS = 2 * c->D * c->D + 1;
    return PICT_OK;
}

/* ------------------------------------------------------------------ */
/* CSR pattern                                                          */
/* ------------------------------------------------------------------ */

/* Row entries of cell (bi,pos): diag + one per non-prescribed face, column
   indices of the neighbours; unsorted (slot order diag,-x,+x,-y,+y,-z,+z;
   -1 for prescribed faces).  PISO_build_matrix :3663-3772. */
static void row_slots(const Ctx *c, int bi, I4 pos, int *idx) {
    const pict_block_desc *b = blk(c, bi);
    for (int k = 0; k < 7; ++k) idx[k] = -1;
    idx[0] = global_index(pos, b);
    for (int face = 0; face < 2 * c->D; ++face) {
        if (at_bound(pos, face, b) && prescribed(face, b)) continue;
        int nb;
        I4 t = neighbor_cell(c, b, pos, face, &nb, bi);
        idx[face + 1] = global_index(t, blk(c, nb));
    }
}
/* findLowestColumnIndex :3479-3489 */
static int lowest(const int *idx) {
    int v = 0x7fffffff, k = -1;
    for (int i = 0; i < 7; ++i) if (idx[i] >= 0 && idx[i] < v) { v = idx[i]; k = i; }
    return k;
}

/* Closed-form row end offset, getCSRMatrixRowEndOffsetFromBlockBoundaries3D :247-312. */
int oracle_csr_row_end_closed_form(const pict_domain_desc *d, int bi, int flatPos, int *row_size) {
    Ctx c; if (init_ctx(&c, d)) return -1;
    const pict_block_desc *b = blk(&c, bi);
    I4 p = unflat(flatPos, b);
    int rs = 2 * c.D + 1, re = (flatPos + 1) * rs;
    if (prescribed(0, b)) { re -= flatPos / b->size[0] + 1; if (p.a[0] == 0) --rs; }
    if (prescribed(1, b)) { re -= (flatPos + 1) / b->size[0]; if (p.a[0] == b->size[0] - 1) --rs; }
    if (c.D > 1) {
        if (prescribed(2, b)) { re -= b->size[0] * p.a[2] + (p.a[1] == 0 ? p.a[0] + 1 : b->size[0]); if (p.a[1] == 0) --rs; }
        if (prescribed(3, b)) { re -= b->size[0] * p.a[2] + (p.a[1] == b->size[1] - 1 ? p.a[0] + 1 : 0); if (p.a[1] == b->size[1] - 1) --rs; }
    }
    if (c.D > 2) {
        if (prescribed(4, b)) { re -= (p.a[2] == 0 ? flatPos + 1 : b->stride[2]); if (p.a[2] == 0) --rs; }
        if (prescribed(5, b)) { re -= (p.a[2] == b->size[2] - 1 ? flatPos + 1 - p.a[2] * b->stride[2] : 0); if (p.a[2] == b->size[2] - 1) --rs; }
    }
    if (row_size) *row_size = rs;
    return re + b->csr_offset;
}

/* Writes csr_row / csr_index by enumeration (diag + neighbours, sorted). */
int oracle_build_csr_pattern(const pict_domain_desc *d) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    int32_t *row = d->csr_row, *ind = d->csr_index;
    int k = 0;
    row[0] = 0;
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        for (int f = 0; f < b->stride[3]; ++f) {
            int idx[7];
            row_slots(&c, bi, unflat(f, b), idx);
            int n = 0; for (int s = 0; s < 7; ++s) n += idx[s] >= 0;
            for (int i = 0; i < n; ++i) { int s = lowest(idx); ind[k++] = idx[s]; idx[s] = -1; }
            row[b->global_offset + f + 1] = k;
        }
    }
    if (k != d->nnz) { snprintf(g_err, sizeof g_err, "oracle: pattern has %d entries, descriptor nnz %d", k, d->nnz); return PICT_ERR_INVALID; }
    return PICT_OK;
}

/* ------------------------------------------------------------------ */
/* forward stages                                                       */
/* ------------------------------------------------------------------ */

static int check_ortho_flags(int flags) {
    if (flags & (PICT_NON_ORTHO_DIRECT_MATRIX | PICT_NON_ORTHO_CENTER_MATRIX | PICT_NON_ORTHO_DIRECT_RHS | PICT_NON_ORTHO_DIAGONAL_RHS)) {
        snprintf(g_err, sizeof g_err, "oracle: non-orthogonal terms not restated");
        return PICT_ERR_UNSUPPORTED;
    }
    return PICT_OK;
}

/* PISO_build_matrix :3601-3865 (velocity, orthogonal terms) */
int oracle_setup_advection_matrix(const pict_domain_desc *d, double time_step, int flags) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    if ((st = check_ortho_flags(flags))) return st;
    const real dt = (real)time_step;
    real *Cv = (real *)d->C_value, *A = (real *)d->A;
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        #pragma omp parallel for schedule(static)
        for (int f = 0; f < b->stride[3]; ++f) {
            I4 pos = unflat(f, b);
            real fl[6];
            fluxes_nd(&c, bi, pos, fl, NULL);
            int idx[7]; real rv[7] = {0, 0, 0, 0, 0, 0, 0};
            for (int k = 0; k < 7; ++k) idx[k] = -1;
            const real det = cell_det(&c, b, pos);
            real diag = det / dt;
            idx[0] = f + b->global_offset;
            const real visc_p = viscosity(&c, b, pos);
            real alpha_p[3] = {1, 1, 1};
            for (int dd = 0; dd < c.D; ++dd) { I4 t = pos; t.a[3] = dd; alpha_p[dd] = laplace_orth(&c, b, t); }
            for (int face = 0; face < 2 * c.D; ++face) {
                int dim = face >> 1;
                real fs = (face & 1) ? 1 : -1;
                if (!(at_bound(pos, face, b) && prescribed(face, b))) {
                    real ff = fs * (real)0.5 * fl[face];
                    diag += ff; rv[face + 1] += ff;
                    int nb;
                    I4 t = neighbor_cell(&c, b, pos, face, &nb, bi);
                    real alpha = laplace_orth(&c, blk(&c, nb), t);
                    t.a[3] = 0;
                    real visc_n = viscosity(&c, blk(&c, nb), t);
                    idx[face + 1] = global_index(t, blk(&c, nb));
                    real vc = (alpha_p[dim] * visc_p + alpha * visc_n) * (real)0.5;
                    diag += vc; rv[face + 1] -= vc;
                } else {
                    int slip = b->bounds[face].velocity_bc == PICT_BC_DIRICHLET ? 0 : 1;
                    diag += (real)((1 - slip) * 2) * visc_p * alpha_p[dim];
                }
            }
            rv[0] = diag;
            int start = d->csr_row[f + b->global_offset];
            int n = 0; for (int s = 0; s < 7; ++s) n += idx[s] >= 0;
            for (int i = 0; i < n; ++i) {
                int s = lowest(idx);
                rv[s] /= det;
                Cv[start + i] = rv[s];
                idx[s] = -1;
            }
            A[f + b->global_offset] = rv[0];
        }
    }
    return PICT_OK;
}

/* Boundary source terms shared by kPISO_build_advection_RHS :4300-4330 and
   PISO_build_pressure_rhs :5166-5188 (FIXED boundaries): sum over faces at the
   block border of  -u_b,d * F_b*n + u_b,d*(1-slip)*nu*2*alpha_b. */
static real fixed_boundary_source(const Ctx *c, const pict_block_desc *b, I4 pos, int d) {
    real s = 0;
    for (int face = 0; face < 2 * c->D; ++face) {
        if (!at_bound(pos, face, b)) continue;
        const pict_boundary_desc *fb = &b->bounds[face];
        if (fb->type != PICT_BOUNDARY_FIXED) continue;
        int ax = face >> 1;
        real normal = (face & 1) ? 1 : -1;
        I4 t = pos; t.a[3] = d; t.a[ax] = 0;
        real vel = fixed_vel(c, fb, t, d);
        t.a[3] = ax;
        real alpha = laplace_orth_fixed(c, fb, t);
        real flux = fixed_contravariant(c, fb, t) * normal;
        s -= vel * flux;
        real visc = viscosity(c, b, pos);
        int slip = fb->velocity_bc == PICT_BC_DIRICHLET ? 0 : 1;
        s += vel * (real)(1 - slip) * visc * 2 * alpha;
    }
    return s;
}

/* getPressureAtWithBounds :2160-2239 (zeroBound=false) */
static real pressure_at(const Ctx *c, const pict_block_desc *b, I4 pos) {
    I4 t = pos; t.a[3] = 0;
    for (int dim = 0; dim < c->D; ++dim) {
        if (pos.a[dim] < 0) {
            const pict_boundary_desc *bd = &b->bounds[2 * dim];
            if (bd->type == PICT_BOUNDARY_FIXED) { t.a[dim] = 0; return ((const real *)b->pressure)[flat4(t, b->stride)]; }
            if (bd->type == PICT_BOUNDARY_CONNECTED) {
                const pict_block_desc *ob = blk(c, bd->connected_block);
                I4 o = connected_pos(c, t, dim, bd, -(pos.a[dim] + 1));
                return ((const real *)ob->pressure)[flat4(o, ob->stride)];
            }
            if (bd->type == PICT_BOUNDARY_PERIODIC) { t.a[dim] += b->size[dim]; return ((const real *)b->pressure)[flat4(t, b->stride)]; }
            return 0;
        }
        if (pos.a[dim] >= b->size[dim]) {
            const pict_boundary_desc *bd = &b->bounds[2 * dim + 1];
            if (bd->type == PICT_BOUNDARY_FIXED) { t.a[dim] = b->size[dim] - 1; return ((const real *)b->pressure)[flat4(t, b->stride)]; }
            if (bd->type == PICT_BOUNDARY_CONNECTED) {
                const pict_block_desc *ob = blk(c, bd->connected_block);
                I4 o = connected_pos(c, t, dim, bd, pos.a[dim] - b->size[dim]);
                return ((const real *)ob->pressure)[flat4(o, ob->stride)];
            }
            if (bd->type == PICT_BOUNDARY_PERIODIC) { t.a[dim] -= b->size[dim]; return ((const real *)b->pressure)[flat4(t, b->stride)]; }
            return 0;
        }
    }
    return ((const real *)b->pressure)[flat4(t, b->stride)];
}

/* getPressureGradient :801-834 */
static void pressure_gradient(const Ctx *c, const pict_block_desc *b, I4 pos, real *g) {
    real gr[3] = {0, 0, 0};
    for (int dim = 0; dim < c->D; ++dim) {
        I4 t = pos; t.a[3] = 0;
        real fac = (real)0.5;
        if (!(pos.a[dim] == 0 && prescribed(2 * dim, b))) t.a[dim] = pos.a[dim] - 1; else fac = 1;
        real vN = pressure_at(c, b, t);
        if (!(pos.a[dim] == b->size[dim] - 1 && prescribed(2 * dim + 1, b))) t.a[dim] = pos.a[dim] + 1;
        else { t.a[dim] = pos.a[dim]; fac = 1; }
        real vP = pressure_at(c, b, t);
        gr[dim] = (vP - vN) * fac;
    }
    if (b->has_transform) {
        const real *T = cell_T(c, b, pos);
        for (int j = 0; j < c->D; ++j) {
            real s = 0;
            for (int i = 0; i < c->D; ++i) s += T_MINV(c, T, i, j) * gr[i];
            g[j] = s;
        }
    } else {
        for (int j = 0; j < c->D; ++j) g[j] = gr[j];
    }
}

/* kPISO_build_advection_RHS :4281-4385 (orthogonal) */
int oracle_setup_advection_velocity(const pict_domain_desc *d, double time_step, int flags, int apply_pressure_gradient) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    if ((st = check_ortho_flags(flags))) return st;
    const real dt = (real)time_step;
    real *rhs = (real *)d->velocity_rhs;
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        #pragma omp parallel for schedule(static)
        for (int f = 0; f < b->stride[3]; ++f) {
            I4 pos = unflat(f, b);
            const real det = cell_det(&c, b, pos);
            real pg[3] = {0, 0, 0};
            if (apply_pressure_gradient) pressure_gradient(&c, b, pos, pg);
            for (int dd = 0; dd < c.D; ++dd) {
                I4 t = pos; t.a[3] = dd;
                real r = det * ((const real *)b->velocity)[flat4(t, b->stride)] / dt;
                r += fixed_boundary_source(&c, b, pos, dd);
                r /= det;
                r += velocity_source(b, t);
                if (apply_pressure_gradient) r -= pg[dd] * dt;
                rhs[(size_t)dd * c.N + b->global_offset + f] = r;
            }
        }
    }
    return PICT_OK;
}

/* PISO_build_pressure_matrix :4797-4963 (orthogonal, no face transform) */
int oracle_setup_pressure_matrix(const pict_domain_desc *d, int flags) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    if ((st = check_ortho_flags(flags))) return st;
    real *Pv = (real *)d->P_value; const real *A = (const real *)d->A;
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        #pragma omp parallel for schedule(static)
        for (int f = 0; f < b->stride[3]; ++f) {
            I4 pos = unflat(f, b);
            int g = b->global_offset + f;
            int idx[7]; real rv[7] = {0, 0, 0, 0, 0, 0, 0};
            for (int k = 0; k < 7; ++k) idx[k] = -1;
            real alphaP[3];
            for (int dd = 0; dd < c.D; ++dd) { I4 t = pos; t.a[3] = dd; alphaP[dd] = laplace_orth(&c, b, t); }
            const real raP = (real)1.0 / A[g];
            idx[0] = g;
            for (int face = 0; face < 2 * c.D; ++face) {
                if (at_bound(pos, face, b) && prescribed(face, b)) continue;
                int dim = face >> 1, nb;
                I4 t = neighbor_cell(&c, b, pos, face, &nb, bi);
                real alphaN = laplace_orth(&c, blk(&c, nb), t);
                t.a[3] = 0;
                int gn = global_index(t, blk(&c, nb));
                real raN = (real)1.0 / A[gn];
                real coef = (real)0.5 * (alphaP[dim] * raP + alphaN * raN);
                rv[0] -= coef; rv[face + 1] += coef; idx[face + 1] = gn;
            }
            int start = d->csr_row[g];
            int n = 0; for (int s = 0; s < 7; ++s) n += idx[s] >= 0;
            for (int i = 0; i < n; ++i) { int s = lowest(idx); Pv[start + i] = rv[s]; idx[s] = -1; }
        }
    }
    return PICT_OK;
}

/* PISO_build_pressure_rhs :5121-5240 */
int oracle_setup_pressure_rhs_only(const pict_domain_desc *d, double time_step) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    const real dt = (real)time_step;
    const real *A = (const real *)d->A, *Cv = (const real *)d->C_value, *vr = (const real *)d->velocity_result;
    real *prhs = (real *)d->pressure_rhs;
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        #pragma omp parallel for schedule(static)
        for (int f = 0; f < b->stride[3]; ++f) {
            I4 pos = unflat(f, b);
            int g = b->global_offset + f;
            const real rDiag = 1 / A[g];
            const real det = cell_det(&c, b, pos);
            int cs = d->csr_row[g], ce = d->csr_row[g + 1];
            for (int dd = 0; dd < c.D; ++dd) {
                I4 t = pos; t.a[3] = dd;
                real velOld = ((const real *)b->velocity)[flat4(t, b->stride)] / dt;
                real H = 0;
                for (int k = cs; k < ce; ++k) {
                    int col = d->csr_index[k];
                    if (col != g) H += Cv[k] * vr[col + (size_t)c.N * dd];
                }
                real S = fixed_boundary_source(&c, b, pos, dd);
                S /= det;
                S += velocity_source(b, t);
                prhs[g + (size_t)c.N * dd] = rDiag * (velOld - H + S);
            }
        }
    }
    return PICT_OK;
}

/* k_computePressureRHSdivergenceFromFlux :5374-5419 */
int oracle_setup_pressure_rhs_div(const pict_domain_desc *d, double time_step, int time_step_norm) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    const real dt = (real)time_step;
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        #pragma omp parallel for schedule(static)
        for (int f = 0; f < b->stride[3]; ++f) {
            real fl[6];
            fluxes_nd(&c, bi, unflat(f, b), fl, (const real *)d->pressure_rhs);
            real div = 0;
            for (int dd = 0; dd < c.D; ++dd) div += fl[2 * dd + 1] - fl[2 * dd];
            if (time_step_norm) div /= dt;
            ((real *)d->pressure_rhs_div)[f + b->global_offset] = div;
        }
    }
    return PICT_OK;
}

/* k_computeVelocityDivergenceFromFlux :6333-6354 */
int oracle_compute_velocity_divergence(const pict_domain_desc *d, real *out) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        for (int f = 0; f < b->stride[3]; ++f) {
            real fl[6];
            fluxes_nd(&c, bi, unflat(f, b), fl, NULL);
            real div = 0;
            for (int dd = 0; dd < c.D; ++dd) div += fl[2 * dd + 1] - fl[2 * dd];
            out[f + b->global_offset] = div;
        }
    }
    return PICT_OK;
}

/* PISO_update_velocity :5947-5980 */
int oracle_correct_velocity(const pict_domain_desc *d, double time_step, int time_step_norm) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    const real dt = (real)time_step;
    const real *A = (const real *)d->A, *prhs = (const real *)d->pressure_rhs;
    real *vr = (real *)d->velocity_result;
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        #pragma omp parallel for schedule(static)
        for (int f = 0; f < b->stride[3]; ++f) {
            I4 pos = unflat(f, b);
            int g = b->global_offset + f;
            real rDiag = 1 / A[g];
            real pg[3];
            pressure_gradient(&c, b, pos, pg);
            for (int dd = 0; dd < c.D; ++dd) {
                real u = -rDiag * pg[dd];
                if (time_step_norm) u *= dt;
                u += prhs[g + (size_t)c.N * dd];
                vr[g + (size_t)c.N * dd] = u;
            }
        }
    }
    return PICT_OK;
}

/* Copy*Result*Blocks :6541-6733 (memcpy per block and component) */
int oracle_copy_velocity_result(const pict_domain_desc *d, int to_blocks) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        for (int dd = 0; dd < c.D; ++dd) {
            real *bv = (real *)b->velocity + (size_t)dd * b->stride[3];
            real *gv = (real *)d->velocity_result + (size_t)dd * c.N + b->global_offset;
            if (to_blocks) memcpy(bv, gv, sizeof(real) * b->stride[3]);
            else memcpy(gv, bv, sizeof(real) * b->stride[3]);
        }
    }
    return PICT_OK;
}
int oracle_copy_pressure_result(const pict_domain_desc *d, int to_blocks) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        real *bp = (real *)b->pressure, *gp = (real *)d->pressure_result + b->global_offset;
        if (to_blocks) memcpy(bp, gp, sizeof(real) * b->stride[3]);
        else memcpy(gp, bp, sizeof(real) * b->stride[3]);
    }
    return PICT_OK;
}

/* ------------------------------------------------------------------ */
/* backward stages (vector-Jacobian products), serial                   */
/* ------------------------------------------------------------------ */

/* Slot -> CSR position of the cell's sorted row (LoadCSRrowNeighborSorted :3503-3563). */
static void slot_positions(const Ctx *c, int bi, I4 pos, int g, int *at) {
    int idx[7];
    row_slots(c, bi, pos, idx);
    for (int k = 0; k < 7; ++k) {
        if (idx[k] < 0) { at[k] = -1; continue; }
        int r = 0;
        for (int j = 0; j < 7; ++j) r += (idx[j] >= 0 && idx[j] < idx[k]);
        at[k] = c->d->csr_row[g] + r;
    }
}

/* scatterViscosityBlock_GRAD :1852-1873 */
static void scatter_viscosity(const Ctx *c, const pict_block_desc *b, I4 p, real g) {
    real *dst = NULL;
    if (!b->viscosity) dst = (real *)c->d->viscosity_grad;
    else if (b->viscosity_static) dst = (real *)b->viscosity_grad;
    else if (b->viscosity_grad) { p.a[3] = 0; dst = (real *)b->viscosity_grad + flat4(p, b->stride); }
    if (dst) *dst += g;
}
/* scatterBlockVelocitySource_GRAD :2255-2264 */
static void scatter_velocity_source(const pict_block_desc *b, I4 p, real g) {
    if (!b->velocity_source_grad) return;
    real *dst = (real *)b->velocity_source_grad;
    if (b->velocity_source_static) dst[p.a[3]] += g; else dst[flat4(p, b->stride)] += g;
}
/* scatterFixedBoundaryData (VELOCITY_GRAD) :2451-2476 */
static void scatter_fixed_velocity(const pict_boundary_desc *fb, I4 p, real g) {
    if (!fb->velocity_grad) return;
    real *dst = (real *)fb->velocity_grad;
    if (fb->velocity_static) dst[p.a[3]] += g; else dst[flat4(p, fb->stride)] += g;
}
/* scatterContravariantComponentGradBoundaryFixed :700-716, :545-560 */
static void scatter_fixed_contravariant(const Ctx *c, const pict_boundary_desc *fb, I4 p, real g) {
    if (!fb->velocity_grad) return;
    int comp = p.a[3];
    if (fb->has_transform) {
        const real *T = bound_T(c, fb, p);
        for (int j = 0; j < c->D; ++j) { p.a[3] = j; scatter_fixed_velocity(fb, p, T_MINV(c, T, comp, j) * T_DET(c, T) * g); }
    } else {
        scatter_fixed_velocity(fb, p, g);
    }
}
/* scatterContravariantComponentGrad :684-698, :527-543 */
static void scatter_contravariant(const Ctx *c, const pict_block_desc *b, I4 p, real g, real *vgg) {
    int comp = p.a[3];
    if (b->has_transform) {
        const real *T = cell_T(c, b, p);
        for (int j = 0; j < c->D; ++j) {
            p.a[3] = j;
            real v = T_MINV(c, T, comp, j) * (T_DET(c, T) * g);
            if (vgg) vgg[(size_t)j * c->N + global_index(p, b)] += v;
            else ((real *)b->velocity_grad)[flat4(p, b->stride)] += v;
        }
    } else {
        if (vgg) vgg[(size_t)comp * c->N + global_index(p, b)] += g;
        else ((real *)b->velocity_grad)[flat4(p, b->stride)] += g;
    }
}
/* ScatterFluxesGradNDLoop :1635-1699 */
static void scatter_fluxes(const Ctx *c, int bi, I4 pos, const real *fg, real *vgg) {
    const pict_block_desc *b = blk(c, bi);
    for (int dim = 0; dim < c->D; ++dim) {
        I4 t = pos; t.a[3] = dim;
        real own = 0;
        for (int up = 0; up < 2; ++up) {
            int face = 2 * dim + up;
            real g = fg[face];
            t.a[dim] = pos.a[dim];
            if (at_bound(pos, face, b)) {
                const pict_boundary_desc *bd = &b->bounds[face];
                if (bd->type == PICT_BOUNDARY_FIXED) {
                    t.a[dim] = 0;
                    scatter_fixed_contravariant(c, bd, t, g);
                } else if (bd->type == PICT_BOUNDARY_CONNECTED) {
                    real of = ((bd->axes[0] & 1) == up) ? (real)-0.5 : (real)0.5;
                    I4 o = connected_pos(c, t, dim, bd, 0);
                    o.a[3] = bd->axes[0] >> 1;
                    scatter_contravariant(c, blk(c, bd->connected_block), o, of * g, vgg);
                    own += (real)0.5 * g;
                } else if (bd->type == PICT_BOUNDARY_PERIODIC) {
                    t.a[dim] = up ? 0 : b->size[dim] - 1;
                    scatter_contravariant(c, b, t, (real)0.5 * g, vgg);
                    own += (real)0.5 * g;
                }
            } else {
                t.a[dim] = pos.a[dim] + (up ? 1 : -1);
                scatter_contravariant(c, b, t, (real)0.5 * g, vgg);
                own += (real)0.5 * g;
            }
        }
        t.a[dim] = pos.a[dim];
        scatter_contravariant(c, b, t, own, vgg);
    }
}
/* VJP of fixed_boundary_source: kPISO_build_advection_RHS_GRAD :4415-4465,
   PISO_build_pressure_rhs_GRAD :5302-5355 (G = gradient of the source term / det) */
static void fixed_boundary_source_grad(const Ctx *c, const pict_block_desc *b, I4 pos, int d, real G) {
    for (int face = 0; face < 2 * c->D; ++face) {
        if (!at_bound(pos, face, b)) continue;
        const pict_boundary_desc *fb = &b->bounds[face];
        if (fb->type != PICT_BOUNDARY_FIXED) continue;
        int ax = face >> 1;
        real normal = (face & 1) ? 1 : -1;
        I4 t = pos; t.a[3] = d; t.a[ax] = 0;
        real vel = fixed_vel(c, fb, t, d);
        t.a[3] = ax;
        real alpha = laplace_orth_fixed(c, fb, t);
        real flux = fixed_contravariant(c, fb, t) * normal;
        real slip = fb->velocity_bc == PICT_BC_DIRICHLET ? 0 : 1;
        real visc = viscosity(c, b, pos);
        scatter_fixed_contravariant(c, fb, t, (-vel * G) * normal);
        t.a[3] = d;
        scatter_fixed_velocity(fb, t, -flux * G + (1 - slip) * visc * 2 * alpha * G);
        scatter_viscosity(c, b, pos, vel * (1 - slip) * 2 * alpha * G);
    }
}

/* PISO_build_matrix_GRAD :3869-4075 */
int oracle_setup_advection_matrix_grad(const pict_domain_desc *d, double time_step, int flags) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    if ((st = check_ortho_flags(flags))) return st;
    (void)time_step;
    const real *Cg = (const real *)d->C_grad, *Ag = (const real *)d->A_grad;
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        for (int f = 0; f < b->stride[3]; ++f) {
            I4 pos = unflat(f, b);
            int g = b->global_offset + f;
            real rdet = 1 / cell_det(&c, b, pos);
            int at[7]; slot_positions(&c, bi, pos, g, at);
            real cg[7];
            for (int k = 0; k < 7; ++k) cg[k] = at[k] >= 0 ? Cg[at[k]] * rdet : 0;
            real diag_grad = Ag[g] * rdet + cg[0];
            real alpha_p[3] = {1, 1, 1};
            for (int dd = 0; dd < c.D; ++dd) { I4 t = pos; t.a[3] = dd; alpha_p[dd] = laplace_orth(&c, b, t); }
            real fg[6] = {0, 0, 0, 0, 0, 0}, visc_grad = 0;
            for (int face = 0; face < 2 * c.D; ++face) {
                int dim = face >> 1;
                real fs = (face & 1) ? 1 : -1;
                if (!(at_bound(pos, face, b) && prescribed(face, b))) {
                    int nb;
                    I4 t = neighbor_cell(&c, b, pos, face, &nb, bi);
                    real alpha = laplace_orth(&c, blk(&c, nb), t);
                    fg[face] = fs * (real)0.5 * (diag_grad + cg[face + 1]);
                    real vcg = diag_grad - cg[face + 1];
                    visc_grad += (real)0.5 * alpha_p[dim] * vcg;
                    t.a[3] = 0;
                    scatter_viscosity(&c, blk(&c, nb), t, (real)0.5 * alpha * vcg);
                } else {
                    real slip = b->bounds[face].velocity_bc == PICT_BC_DIRICHLET ? 0 : 1;
                    visc_grad += (1 - slip) * 2 * alpha_p[dim] * diag_grad;
                }
            }
            if (b->velocity_grad) scatter_fluxes(&c, bi, pos, fg, NULL);
            scatter_viscosity(&c, b, pos, visc_grad);
        }
    }
    return PICT_OK;
}

/* kPISO_build_advection_RHS_GRAD :4388-4476 */
int oracle_setup_advection_velocity_grad(const pict_domain_desc *d, double time_step, int flags) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    if ((st = check_ortho_flags(flags))) return st;
    const real dt = (real)time_step;
    const real *rg = (const real *)d->velocity_rhs_grad;
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        for (int f = 0; f < b->stride[3]; ++f) {
            I4 pos = unflat(f, b);
            real det = cell_det(&c, b, pos);
            for (int dd = 0; dd < c.D; ++dd) {
                I4 t = pos; t.a[3] = dd;
                real G = rg[(size_t)dd * c.N + b->global_offset + f];
                scatter_velocity_source(b, t, G);
                G /= det;
                fixed_boundary_source_grad(&c, b, pos, dd, G);
                G *= det;
                if (b->velocity_grad) ((real *)b->velocity_grad)[flat4(t, b->stride)] = G / dt;
            }
        }
    }
    return PICT_OK;
}

/* PISO_build_pressure_matrix_GRAD :4967-5115 */
int oracle_setup_pressure_matrix_grad(const pict_domain_desc *d, int flags) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    if ((st = check_ortho_flags(flags))) return st;
    const real *Pg = (const real *)d->P_grad, *A = (const real *)d->A;
    real *Ag = (real *)d->A_grad;
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        for (int f = 0; f < b->stride[3]; ++f) {
            I4 pos = unflat(f, b);
            int g = b->global_offset + f;
            int at[7]; slot_positions(&c, bi, pos, g, at);
            real raP_grad = 0;
            for (int face = 0; face < 2 * c.D; ++face) {
                if (at_bound(pos, face, b) && prescribed(face, b)) continue;
                int dim = face >> 1, nb;
                I4 t = neighbor_cell(&c, b, pos, face, &nb, bi);
                real alphaN = laplace_orth(&c, blk(&c, nb), t);
                I4 tp = pos; tp.a[3] = dim;
                real alphaP = laplace_orth(&c, b, tp);
                t.a[3] = 0;
                int gn = global_index(t, blk(&c, nb));
                real raN = 1 / A[gn];
                real cg = Pg[at[face + 1]] - Pg[at[0]];
                raP_grad += (real)0.5 * alphaP * cg;
                Ag[gn] += -((real)0.5 * alphaN * cg) * raN * raN;
            }
            real raP = 1 / A[g];
            Ag[g] += -raP_grad * raP * raP;
        }
    }
    return PICT_OK;
}

/* k_computePressureRHSdivergenceFromFlux_GRAD :5423-5452 */
int oracle_setup_pressure_rhs_div_grad(const pict_domain_desc *d, double time_step, int time_step_norm) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    const real dt = (real)time_step;
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        for (int f = 0; f < b->stride[3]; ++f) {
            real dg = ((const real *)d->pressure_rhs_div_grad)[b->global_offset + f];
            if (time_step_norm) dg /= dt;
            real fg[6];
            for (int dd = 0; dd < c.D; ++dd) { fg[2 * dd + 1] = dg; fg[2 * dd] = -dg; }
            scatter_fluxes(&c, bi, unflat(f, b), fg, (real *)d->pressure_rhs_grad);
        }
    }
    return PICT_OK;
}

/* PISO_build_pressure_rhs_GRAD :5243-5369 */
int oracle_setup_pressure_rhs_only_grad(const pict_domain_desc *d, double time_step) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    const real dt = (real)time_step;
    const real *A = (const real *)d->A, *Cv = (const real *)d->C_value, *vr = (const real *)d->velocity_result;
    const real *prhs = (const real *)d->pressure_rhs, *pg = (const real *)d->pressure_rhs_grad;
    real *vrg = (real *)d->velocity_result_grad, *Cg = (real *)d->C_grad, *Ag = (real *)d->A_grad;
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        for (int f = 0; f < b->stride[3]; ++f) {
            I4 pos = unflat(f, b);
            int g = b->global_offset + f;
            real rDiag = 1 / A[g], det = cell_det(&c, b, pos), rDiag_grad = 0;
            int cs = d->csr_row[g], ce = d->csr_row[g + 1];
            real cvg[7] = {0, 0, 0, 0, 0, 0, 0};
            for (int dd = 0; dd < c.D; ++dd) {
                I4 t = pos; t.a[3] = dd;
                size_t o = (size_t)c.N * dd;
                real pGrad = pg[g + o];
                rDiag_grad += prhs[g + o] / rDiag * pGrad;
                if (b->velocity_grad) ((real *)b->velocity_grad)[flat4(t, b->stride)] = rDiag / dt * pGrad;
                real w = rDiag * pGrad;
                for (int k = cs; k < ce; ++k) {
                    int col = d->csr_index[k];
                    if (col == g) continue;
                    if (vrg) vrg[col + o] += -Cv[k] * w;
                    cvg[k - cs] -= vr[col + o] * w;
                }
                real S_grad = w;
                scatter_velocity_source(b, t, S_grad);
                S_grad /= det;
                fixed_boundary_source_grad(&c, b, pos, dd, S_grad);
            }
            if (Cg) for (int k = cs; k < ce; ++k) Cg[k] = cvg[k - cs];
            if (Ag) Ag[g] += rDiag_grad * (-rDiag * rDiag);
        }
    }
    return PICT_OK;
}

/* scatterPressureGradToWithBounds :3411-3477 */
static void scatter_pressure_to(const Ctx *c, const pict_block_desc *b, I4 pos, real g) {
    I4 t = pos; t.a[3] = 0;
    for (int dim = 0; dim < c->D; ++dim) {
        if (pos.a[dim] < 0 || pos.a[dim] >= b->size[dim]) {
            int lower = pos.a[dim] < 0;
            const pict_boundary_desc *bd = &b->bounds[2 * dim + (lower ? 0 : 1)];
            if (bd->type == PICT_BOUNDARY_FIXED) {
                t.a[dim] = lower ? 0 : b->size[dim] - 1;
                ((real *)b->pressure_grad)[flat4(t, b->stride)] += g;
            } else if (bd->type == PICT_BOUNDARY_CONNECTED) {
                const pict_block_desc *ob = blk(c, bd->connected_block);
                I4 o = connected_pos(c, t, dim, bd, lower ? -(pos.a[dim] + 1) : pos.a[dim] - b->size[dim]);
                ((real *)ob->pressure_grad)[flat4(o, ob->stride)] += g;
            } else if (bd->type == PICT_BOUNDARY_PERIODIC) {
                t.a[dim] += lower ? b->size[dim] : -b->size[dim];
                ((real *)b->pressure_grad)[flat4(t, b->stride)] += g;
            }
            return;
        }
    }
    ((real *)b->pressure_grad)[flat4(t, b->stride)] += g;
}

/* PISO_update_velocity_GRAD :6250-6293 with scatterPressureGradientGrad :838-870 */
int oracle_correct_velocity_grad(const pict_domain_desc *d, double time_step, int time_step_norm) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    const real dt = (real)time_step;
    const real *A = (const real *)d->A, *vrg = (const real *)d->velocity_result_grad;
    real *prg = (real *)d->pressure_rhs_grad, *Ag = (real *)d->A_grad;
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        for (int f = 0; f < b->stride[3]; ++f) {
            I4 pos = unflat(f, b);
            int g = b->global_offset + f;
            real rDiag = 1 / A[g], pg[3], pgg[3] = {0, 0, 0}, rDiag_grad = 0;
            pressure_gradient(&c, b, pos, pg);
            for (int dd = 0; dd < c.D; ++dd) {
                size_t o = (size_t)c.N * dd;
                real vug = vrg[g + o];
                prg[g + o] = vug;
                if (time_step_norm) vug *= dt;
                pgg[dd] = -rDiag * vug;
                rDiag_grad -= pg[dd] * vug;
            }
            if (b->pressure_grad) {
                real v[3];
                if (b->has_transform) {
                    const real *T = cell_T(&c, b, pos);
                    for (int i = 0; i < c.D; ++i) { real s = 0; for (int j = 0; j < c.D; ++j) s += T_MINV(&c, T, i, j) * pgg[j]; v[i] = s; }
                } else {
                    for (int i = 0; i < c.D; ++i) v[i] = pgg[i];
                }
                for (int dim = 0; dim < c.D; ++dim) {
                    I4 pP = pos, pN = pos;
                    real fac = (real)0.5;
                    if (!(pos.a[dim] == 0 && prescribed(2 * dim, b))) pN.a[dim] = pos.a[dim] - 1; else fac = 1;
                    if (!(pos.a[dim] == b->size[dim] - 1 && prescribed(2 * dim + 1, b))) pP.a[dim] = pos.a[dim] + 1; else fac = 1;
                    scatter_pressure_to(&c, b, pN, v[dim] * -fac);
                    scatter_pressure_to(&c, b, pP, v[dim] * fac);
                }
            }
            Ag[g] = -rDiag_grad * rDiag * rDiag;
        }
    }
    return PICT_OK;
}

/* Copy{Velocity,Pressure}ResultGrad{To,From}Blocks :6819-6915 */
int oracle_copy_velocity_result_grad(const pict_domain_desc *d, int to_blocks) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        for (int dd = 0; dd < c.D; ++dd) {
            real *bv = (real *)b->velocity_grad + (size_t)dd * b->stride[3];
            real *gv = (real *)d->velocity_result_grad + (size_t)dd * c.N + b->global_offset;
            if (to_blocks) memcpy(bv, gv, sizeof(real) * b->stride[3]);
            else memcpy(gv, bv, sizeof(real) * b->stride[3]);
        }
    }
    return PICT_OK;
}
int oracle_copy_pressure_result_grad_from_blocks(const pict_domain_desc *d) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        memcpy((real *)d->pressure_result_grad + b->global_offset, b->pressure_grad, sizeof(real) * b->stride[3]);
    }
    return PICT_OK;
}

/* ------------------------------------------------------------------ */
/* Krylov solvers                                                       */
/* ------------------------------------------------------------------ */

typedef struct { const int32_t *row, *col; const real *val; int n; int transpose; } Mat;

/* y = A x or A^T x (cusparseSpMV, CSR) */
static void spmv(const Mat *m, const real *x, real *y) {
    if (!m->transpose) {
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < m->n; ++i) {
            real s = 0;
            for (int k = m->row[i]; k < m->row[i + 1]; ++k) s += m->val[k] * x[m->col[k]];
            y[i] = s;
        }
    } else {
        for (int i = 0; i < m->n; ++i) y[i] = 0;
        for (int i = 0; i < m->n; ++i)
            for (int k = m->row[i]; k < m->row[i + 1]; ++k) y[m->col[k]] += m->val[k] * x[i];
    }
}
/* FP64 dot product in a fixed order whatever the thread count: 4096-element blocks summed
   sequentially, then the block sums in block order (an OpenMP reduction combines the threads'
   sums in an unspecified order, which makes iteration counts vary from run to run). */
static double dotd(const real *a, const real *b, int n) {
    enum { BLK = 4096 };
    const int nb = (n + BLK - 1) / BLK;
    double stack_sums[64];
    double *bs = nb <= 64 ? stack_sums : (double *)malloc(sizeof(double) * (size_t)nb);
    if (!bs) return NAN;
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < nb; ++k) {
        const int e = (k + 1) * BLK < n ? (k + 1) * BLK : n;
        double t = 0;
        for (int i = k * BLK; i < e; ++i) t += (double)a[i] * (double)b[i];
        bs[k] = t;
    }
    double s = 0;
    for (int k = 0; k < nb; ++k) s += bs[k];
    if (bs != stack_sums) free(bs);
    return s;
}
/* ComputeConvergenceCriterion, cg_solver_kernel.cu:77-110 */
static double criterion(const real *r, int n, int conv) {
    double s = 0, m = 0;
    switch (conv) {
    case PICT_CONV_NORM2: return sqrt(dotd(r, r, n));
    case PICT_CONV_NORM2_NORMALIZED: return sqrt(dotd(r, r, n)) / sqrt((double)n);
    case PICT_CONV_ABS_SUM: for (int i = 0; i < n; ++i) s += fabs((double)r[i]); return s;
    case PICT_CONV_ABS_MEAN: for (int i = 0; i < n; ++i) s += fabs((double)r[i]); return s / n;
    case PICT_CONV_ABS_MAX: for (int i = 0; i < n; ++i) if (fabs((double)r[i]) > m) m = fabs((double)r[i]); return m;
    }
    return NAN;
}

/* cgSolveGPU, cg_solver_kernel.cu:114-456 (without rank-deficient option) */
static void cg_one(const Mat *m, const real *f, real *x, const pict_solve_opts *o, pict_solve_info *info, real *w) {
    int n = m->n;
    real *r = w, *rw = w + n, *p = w + 2 * n;
    real *best = o->return_best_result ? w + 3 * n : NULL;
    info->final_residual = 0; info->used_iterations = -1; info->converged = 0; info->is_finite_residual = 1;
    double rho, rhop, alpha, beta, bestc = 0, lastc = 0; int bestit = -1, rising = 0;
    spmv(m, x, r);
    for (int i = 0; i < n; ++i) { r[i] = f[i] - r[i]; p[i] = r[i]; }
    rho = dotd(r, r, n);
    for (int it = 0; it < o->max_iterations; ++it) {
        if (o->residual_reset_steps > 0 && (it + 1) % o->residual_reset_steps == 0) {
            spmv(m, x, r);
            for (int i = 0; i < n; ++i) { r[i] = f[i] - r[i]; p[i] = r[i]; }
            rho = dotd(r, r, n);
        }
        spmv(m, p, rw);
        alpha = rho / dotd(p, rw, n);
        for (int i = 0; i < n; ++i) { x[i] += (real)alpha * p[i]; r[i] -= (real)alpha * rw[i]; }
        double crit = criterion(r, n, o->convergence_criterion);
        if (!isfinite(crit)) { info->converged = 0; info->used_iterations = it; info->final_residual = crit; info->is_finite_residual = 0; break; }
        if (best) {
            if (it == 0 || crit < bestc) { bestc = crit; bestit = it; memcpy(best, x, sizeof(real) * n); }
            if (it > 0 && crit >= lastc) ++rising; else rising = 0;
            lastc = crit;
        }
        info->used_iterations = it; info->final_residual = crit;
        if (crit < o->tolerance) { info->converged = 1; info->is_finite_residual = 1; break; }
        if (best && (it == o->max_iterations - 1 || rising >= 100)) {
            memcpy(x, best, sizeof(real) * n);
            info->converged = 0; info->used_iterations = bestit; info->final_residual = bestc; info->is_finite_residual = 1;
            break;
        }
        rhop = rho;
        rho = dotd(r, r, n);
        beta = rho / rhop;
        for (int i = 0; i < n; ++i) p[i] = r[i] + (real)beta * p[i];
    }
}

/* bicgstabSolveGPU without preconditioner, bicgstab_solver_kernel.cu:222-365 */
static void bicg_one(const Mat *m, const real *f, real *x, const pict_solve_opts *o, pict_solve_info *info, real *w) {
    int n = m->n;
    real *r = w, *rw = w + n, *p = w + 2 * n, *v = w + 3 * n, *s = w + 4 * n, *t = w + 5 * n;
    info->final_residual = 0; info->used_iterations = -1; info->converged = 0; info->is_finite_residual = 1;
    double rho = 0, rhop, alpha = 0, omega = 0, beta;
    spmv(m, x, r);
    for (int i = 0; i < n; ++i) { r[i] = f[i] - r[i]; rw[i] = r[i]; p[i] = r[i]; v[i] = 0; }
    double nrmr0 = criterion(r, n, o->convergence_criterion);
    if (nrmr0 < o->tolerance) { info->converged = 1; info->used_iterations = -1; info->final_residual = nrmr0; return; }
    for (int it = 0; it < o->max_iterations; ++it) {
        rhop = rho;
        rho = dotd(rw, r, n);
        if (it > 0) {
            beta = (rho / rhop) * (alpha / omega);
            for (int i = 0; i < n; ++i) p[i] = r[i] + (real)beta * (p[i] - (real)omega * v[i]);
        }
        spmv(m, p, v);
        alpha = rho / dotd(rw, v, n);
        for (int i = 0; i < n; ++i) { r[i] -= (real)alpha * v[i]; x[i] += (real)alpha * p[i]; }
        double nr = criterion(r, n, o->convergence_criterion);
        if (!isfinite(nr)) { info->converged = 0; info->used_iterations = it; info->final_residual = nr; info->is_finite_residual = 0; break; }
        info->used_iterations = it; info->final_residual = nr;
        if (nr < o->tolerance) { info->converged = 1; break; }
        memcpy(s, r, sizeof(real) * n);
        spmv(m, s, t);
        omega = dotd(t, r, n) / dotd(t, t, n);
        for (int i = 0; i < n; ++i) { x[i] += (real)omega * s[i]; r[i] -= (real)omega * t[i]; }
        nr = criterion(r, n, o->convergence_criterion);
        info->final_residual = nr;
        if (nr < o->tolerance) { info->converged = 1; info->used_iterations = it + 1; break; }
    }
}

int oracle_solve_linear(const pict_csr_desc *A, const real *rhs, real *x, int nrhs, const pict_solve_opts *o, pict_solve_info *infos) {
    if (A->dtype != ORACLE_DTYPE) { snprintf(g_err, sizeof g_err, "oracle: csr dtype mismatch"); return PICT_ERR_INVALID; }
    if (o->matrix_rank_deficient || (o->use_bicg && o->bicg_with_preconditioner)) {
        snprintf(g_err, sizeof g_err, "oracle: option not restated"); return PICT_ERR_UNSUPPORTED;
    }
    Mat m = {A->row, A->index, (const real *)A->value, A->n, o->transpose_a};
    real *w = (real *)malloc(sizeof(real) * (size_t)A->n * 6);
    for (int k = 0; k < nrhs; ++k) {
        const real *f = rhs + (size_t)k * A->n;
        real *xk = x + (size_t)k * A->n;
        if (o->use_bicg) bicg_one(&m, f, xk, o, infos + k, w);
        else cg_one(&m, f, xk, o, infos + k, w);
    }
    free(w);
    return PICT_OK;
}

/* OuterProductToSparseMatrix (SDDMM), cg_solver_kernel.cu:869-908 */
int oracle_sparse_outer_product(const real *a, const real *b, const pict_csr_desc *pat, real *out) {
    for (int i = 0; i < pat->n; ++i)
        for (int k = pat->row[i]; k < pat->row[i + 1]; ++k) out[k] = a[i] * b[pat->index[k]];
    return PICT_OK;
}

/* ------------------------------------------------------------------ */
/* metrics from vertex coordinates, grid_gen.cu                         */
/* ------------------------------------------------------------------ */

static real det_n(const real *M, int D) {
    if (D == 1) return M[0];
    if (D == 2) return M[0] * M[3] - M[1] * M[2];
    return M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) + M[2] * (M[3] * M[7] - M[4] * M[6]);
}
/* inverse via adjugate / determinant (transformations.h:476-560) */
static void inv_n(const real *M, real det, int D, real *R) {
    if (D == 1) { R[0] = 1 / M[0]; return; }
    if (D == 2) { R[0] = M[3] / det; R[1] = -M[1] / det; R[2] = -M[2] / det; R[3] = M[0] / det; return; }
    R[0] = (M[4] * M[8] - M[5] * M[7]) / det; R[1] = (M[2] * M[7] - M[1] * M[8]) / det; R[2] = (M[1] * M[5] - M[2] * M[4]) / det;
    R[3] = (M[5] * M[6] - M[3] * M[8]) / det; R[4] = (M[0] * M[8] - M[2] * M[6]) / det; R[5] = (M[2] * M[3] - M[0] * M[5]) / det;
    R[6] = (M[3] * M[7] - M[4] * M[6]) / det; R[7] = (M[1] * M[6] - M[0] * M[7]) / det; R[8] = (M[0] * M[4] - M[1] * M[3]) / det;
}

/* k_CoordsToTransforms grid_gen.cu:284-339 */
int oracle_coords_to_transforms(const real *coords, int D, const int32_t vs[3], real *out) {
    int VX = vs[0], VY = D > 1 ? vs[1] : 1, VZ = D > 2 ? vs[2] : 1;
    int CX = VX - 1, CY = D > 1 ? VY - 1 : 1, CZ = D > 2 ? VZ - 1 : 1;
    int TS = 2 * D * D + 1, nv = 1 << D;
    size_t vstride_c = (size_t)VX * VY * VZ;
    for (int z = 0; z < CZ; ++z) for (int y = 0; y < CY; ++y) for (int x = 0; x < CX; ++x) {
        real fc[6][3];
        memset(fc, 0, sizeof fc);
        for (int v = 0; v < nv; ++v) {
            int vx = x + (v & 1), vy = y + ((v & 2) != 0), vz = z + ((v & 4) != 0);
            real co[3];
            for (int k = 0; k < D; ++k) co[k] = coords[k * vstride_c + vx + (size_t)VX * vy + (size_t)VX * VY * vz];
            for (int f = 0; f < D; ++f) { int up = (v & (1 << f)) != 0; for (int k = 0; k < D; ++k) fc[f * 2 + up][k] += co[k]; }
        }
        real norm = (real)1.0 / (real)(1 << (D - 1));
        for (int f = 0; f < 2 * D; ++f) for (int k = 0; k < D; ++k) fc[f][k] *= norm;
        real *T = out + ((size_t)x + (size_t)CX * y + (size_t)CX * CY * z) * TS;
        for (int i = 0; i < D; ++i) for (int k = 0; k < D; ++k) T[i * D + k] = fc[k * 2 + 1][i] - fc[k * 2][i];
        real det = det_n(T, D);
        inv_n(T, det, D, T + D * D);
        T[2 * D * D] = det;
    }
    return PICT_OK;
}

/* k_CoordsToFaceTransforms grid_gen.cu:383-479 */
int oracle_coords_to_face_transforms(const real *coords, int D, const int32_t vs[3], real *out) {
    int SX = vs[0], SY = D > 1 ? vs[1] : 1, SZ = D > 2 ? vs[2] : 1;
    int size[3] = {SX, SY, SZ};
    int TS = 2 * D * D + 1;
    size_t cstride = (size_t)SX * SY * SZ;
    int nfv = 1 << (D - 1), nev = D >= 2 ? 1 << (D - 2) : 1;
    for (int z = 0; z < SZ; ++z) for (int y = 0; y < SY; ++y) for (int x = 0; x < SX; ++x) {
        int pos[3] = {x, y, z};
        for (int dim = 0; dim < D; ++dim) {
            int skip = 0;
            for (int i = 1; i < D; ++i) { int td = (dim + i) % D; if (pos[td] == size[td] - 1) skip = 1; }
            if (skip) continue;
            real t[3][3];
            memset(t, 0, sizeof t);
            real nw = (real)0.5;
            for (int up = 0; up < 2; ++up) {
                int fs = up * 2 - 1;
                int atb = up ? pos[dim] == size[dim] - 1 : pos[dim] == 0;
                int fp[3] = {x, y, z};
                if (!atb) fp[dim] += fs; else nw = 1;
                for (int v = 0; v < nfv; ++v) {
                    int vp[3] = {fp[0], fp[1], fp[2]};
                    for (int i = 1; i < D; ++i) vp[(dim + i) % D] += (v & i) != 0;
                    for (int k = 0; k < D; ++k) t[k][dim] += coords[k * cstride + vp[0] + (size_t)SX * vp[1] + (size_t)SX * SY * vp[2]] * fs;
                }
            }
            nw /= (real)nfv;
            for (int k = 0; k < D; ++k) t[k][dim] *= nw;
            real tn = (real)1.0 / (real)nev;
            for (int kk = 1; kk < D; ++kk) {
                int td = (dim + kk) % D;
                for (int v = 0; v < nfv; ++v) {
                    int vp[3] = {x, y, z};
                    for (int i = 1; i < D; ++i) vp[(dim + i) % D] += (v & i) != 0;
                    int tup = (v & kk) != 0;
                    int tfs = tup ? 1 : -1;
                    for (int k = 0; k < D; ++k) t[k][td] += coords[k * cstride + vp[0] + (size_t)SX * vp[1] + (size_t)SX * SY * vp[2]] * tfs;
                }
                for (int k = 0; k < D; ++k) t[k][td] *= tn;
            }
            real *T = out + ((size_t)dim * cstride + x + (size_t)SX * y + (size_t)SX * SY * z) * TS;
            for (int i = 0; i < D; ++i) for (int k = 0; k < D; ++k) T[i * D + k] = t[i][k];
            real det = det_n(T, D);
            inv_n(T, det, D, T + D * D);
            T[2 * D * D] = det;
        }
    }
    return PICT_OK;
}
