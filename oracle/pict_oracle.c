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

/* getFaceTransformPtr :874-889: the transform of one face of a cell, from the block's face transforms laid out
   [D][(z+1)][(y+1)][x+1] (one more vertex along every spatial axis; the face of `face` sits at the cell's index,
   one further along its axis for an upper face) */
static const real *face_T(const Ctx *c, const pict_block_desc *b, I4 pos, int face) {
    int dim = face >> 1, fs[3];
    pos.a[dim] += face & 1;
    for (int i = 0; i < 3; ++i) fs[i] = b->size[i] + (i < c->D ? 1 : 0);
    size_t idx = (size_t)pos.a[0] + (size_t)fs[0] * ((size_t)pos.a[1] + (size_t)fs[1] * ((size_t)pos.a[2] + (size_t)fs[2] * dim));
    return (const real *)b->face_transform + idx * c->TS;
}


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
/* getLaplaceCoefficientOrthogonalFace :1226-1241: det |Minv[axis]|^2 of the face's own transform (1 without face
   transforms) */
static real laplace_orth_face(const Ctx *c, const pict_block_desc *b, I4 p, int face) {
    if (!b->has_face_transform) return 1;
    const real *T = face_T(c, b, p, face);
    const int ax = face >> 1;
    real s = 0;
    for (int j = 0; j < c->D; ++j) s += T_MINV(c, T, ax, j) * T_MINV(c, T, ax, j);
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


/* ------------------------------------------------------------------ */
/* non-orthogonal terms (flags of PISO_multiblock_cuda.h:9-13)          */
/* ------------------------------------------------------------------ */

/* Data a corner or neighbour value is read from: the GridDataType values the reference's
   stages pass (domain_structs_gpu.h:28-65).  Gradient variants scatter into the matching
   gradient tensors. */
enum { GD_VELOCITY = 0, GD_VELOCITY_RESULT = 1, GD_PRESSURE_RESULT = 2, GD_IS_FIXED = 3,
       GD_SCALAR = 4, GD_SCALAR_RESULT = 5 /* passive scalar: p.w is the channel */ };
#define GD_GRAD 8
#define GD_BASE(t) ((t) & 7)
static inline int gd_is_pressure(int t) { return GD_BASE(t) == GD_PRESSURE_RESULT; }
static inline int gd_is_scalar(int t) { return GD_BASE(t) == GD_SCALAR || GD_BASE(t) == GD_SCALAR_RESULT; }

/* getViscosity(forPassiveScalar) :1788-1800: the scalar viscosity of channel ch, or the domain's */
static real scalar_viscosity(const Ctx *c, int ch) {
    const real *v = (const real *)c->d->scalar_viscosity;
    if (!v) return (real)c->d->viscosity;
    return v[c->d->scalar_viscosity_static ? 0 : ch];
}
/* FixedBoundary passive scalar of channel ch (getFixedBoundaryData :2419-2437), p on the boundary grid */
static real fixed_scalar(const pict_boundary_desc *fb, I4 p, int ch) {
    if (fb->scalar_static) return ((const real *)fb->scalar)[ch];
    p.a[3] = ch;
    return ((const real *)fb->scalar)[flat4(p, fb->stride)];
}
/* getFixedBoundaryType (PASSIVE_SCALAR) :2360-2370 */
static int fixed_scalar_bc(const pict_boundary_desc *fb, int ch) {
    return ((fb->scalar_bc_mask >> (ch & 31)) & 1) ? PICT_BC_NEUMANN : PICT_BC_DIRICHLET;
}

static inline int face_sign(int face) { return (face & 1) ? 1 : -1; }
static inline int i4_eq(const I4 a, const I4 b) {
    return a.a[0] == b.a[0] && a.a[1] == b.a[1] && a.a[2] == b.a[2] && a.a[3] == b.a[3];
}

/* getLaplaceCoefficientCenterIndex :1253-1255 */
static inline int lap_ci(int c1, int c2) { return c1 + c2 + ((c1 >> 1) | (c2 >> 1)); }

/* dot of Minv rows r1, r2 (transformations.h dot: ascending sum from 0) */
static inline real minv_dot(const Ctx *c, const real *T, int r1, int r2) {
    real s = 0;
    for (int j = 0; j < c->D; ++j) s += T_MINV(c, T, r1, j) * T_MINV(c, T, r2, j);
    return s;
}

/* getLaplaceCoefficientsFull :1261-1287: out[lap_ci(c1,c2)] = det Minv_c1 . Minv_c2 */
static void lap_full(const Ctx *c, const pict_block_desc *b, I4 p, real *out) {
    for (int c1 = 0; c1 < c->D; ++c1)
        for (int c2 = 0; c2 <= c1; ++c2) {
            if (b->has_transform) {
                const real *T = cell_T(c, b, p);
                out[lap_ci(c1, c2)] = T_DET(c, T) * minv_dot(c, T, c1, c2);
            } else {
                out[lap_ci(c1, c2)] = c1 == c2 ? 1 : 0;
            }
        }
}
/* getLaplaceCoefficientsSingleAxis :1289-1313 */
static void lap_single_axis(const Ctx *c, const pict_block_desc *b, I4 p, int axis, real *out) {
    for (int c1 = 0; c1 < c->D; ++c1) {
        if (b->has_transform) {
            const real *T = cell_T(c, b, p);
            out[c1] = T_DET(c, T) * minv_dot(c, T, c1, axis);
        } else {
            out[c1] = c1 == axis ? 1 : 0;
        }
    }
}
/* getLaplaceCoefficientsNeighbourBoundary :1468-1490 (FixedBoundary) */
static void lap_neighbour_boundary(const Ctx *c, const pict_boundary_desc *fb, I4 p, int axis, real *out) {
    for (int c1 = 0; c1 < c->D; ++c1) {
        if (fb->has_transform) {
            I4 q = p; q.a[3] = 0; q.a[axis] = 0;
            const real *T = bound_T(c, fb, q);
            out[c1] = T_DET(c, T) * minv_dot(c, T, c1, axis);
        } else {
            out[c1] = c1 == axis ? 1 : 0;
        }
    }
}

/* computeConnectedDir :314-323 */
static inline int connected_dir(const Ctx *c, int dir, int bdim, const pict_boundary_desc *cb) {
    int rel = posmod((dir >> 1) - bdim, c->D);
    return cb->axes[rel] ^ (dir & 1);
}

/* A resolved neighbour cell (resolveNeighborCell :431-476) with the axis mapping of the connection. */
typedef struct { int bi; I4 pos; int map[3]; } NCell;
static NCell resolve_neighbor(const Ctx *c, int bi, I4 pos, int dir) {
    const pict_block_desc *b = blk(c, bi);
    NCell n; n.bi = bi; n.pos = pos; n.map[0] = 0; n.map[1] = 2; n.map[2] = 4;
    int axis = dir >> 1;
    if (at_bound(pos, dir, b)) {
        const pict_boundary_desc *bd = &b->bounds[dir];
        if (bd->type == PICT_BOUNDARY_CONNECTED) {
            n.bi = bd->connected_block;
            n.pos = connected_pos(c, pos, axis, bd, 0);
            for (int dd = 0; dd < c->D; ++dd) n.map[dd] = connected_dir(c, n.map[dd], axis, bd);
        } else if (bd->type == PICT_BOUNDARY_PERIODIC) {
            n.pos.a[axis] = (dir & 1) ? 0 : b->size[axis] - 1;
        }
    } else {
        n.pos.a[axis] += face_sign(dir);
    }
    return n;
}
/* getLaplaceCoefficientsSingleAxisNeighbor(NeighborCellInfo) :1415-1452 (block cells only) */
static void lap_single_axis_neighbor(const Ctx *c, const NCell *n, int dir, real *out) {
    int axis = dir >> 1;
    const pict_block_desc *nb = blk(c, n->bi);
    for (int c1 = 0; c1 < c->D; ++c1) out[c1] = 0;
    if (nb->has_transform) {
        I4 q = n->pos; q.a[3] = 0;
        const real *T = cell_T(c, nb, q);
        int am = n->map[axis] >> 1;
        for (int c1 = 0; c1 < c->D; ++c1) out[c1] = T_DET(c, T) * minv_dot(c, T, n->map[c1] >> 1, am);
    } else {
        out[axis] = 1;
    }
}

/* getBlockData :2320-2331 through getBlockCellDataPtr :2305-2317 */
static real block_data(const Ctx *c, int bi, I4 p, int type) {
    const pict_block_desc *b = blk(c, bi);
    switch (GD_BASE(type)) {
        case GD_VELOCITY: return ((const real *)b->velocity)[flat4(p, b->stride)];
        case GD_VELOCITY_RESULT: return ((const real *)c->d->velocity_result)[(size_t)p.a[3] * c->N + global_index(p, b)];
        case GD_PRESSURE_RESULT: return ((const real *)c->d->pressure_result)[(size_t)p.a[3] * c->N + global_index(p, b)];
        case GD_SCALAR: return ((const real *)b->scalar)[flat4(p, b->stride)];
        case GD_SCALAR_RESULT: return ((const real *)c->d->scalar_result)[(size_t)p.a[3] * c->N + global_index(p, b)];
        default: return 0;
    }
}
/* getFixedBoundaryData :2376-2446 (FIXED; the pressure of a FixedBoundary is a static zero) */
static real fixed_data(const Ctx *c, const pict_block_desc *b, I4 p, int face, int type) {
    if (GD_BASE(type) == GD_IS_FIXED) return 1;
    if (type & GD_GRAD) return 0;
    const pict_boundary_desc *fb = &b->bounds[face];
    if (fb->type != PICT_BOUNDARY_FIXED) return 0;
    if (GD_BASE(type) == GD_PRESSURE_RESULT) return 0;
    p.a[face >> 1] = 0;
    if (gd_is_scalar(type)) return fixed_scalar(fb, p, p.a[3]);
    if (fb->velocity_static) return ((const real *)fb->velocity)[p.a[3]];
    return ((const real *)fb->velocity)[flat4(p, fb->stride)];
}
/* getFixedBoundaryType :2351-2374 (velocity: the boundary's type; pressure: Neumann; passive scalar: the type
   of channel ch, pos.w in the reference) */
static int fixed_type(const pict_block_desc *b, int face, int type, int ch) {
    const pict_boundary_desc *fb = &b->bounds[face];
    if (fb->type != PICT_BOUNDARY_FIXED) return PICT_BC_DIRICHLET;
    switch (GD_BASE(type)) {
        case GD_VELOCITY:
        case GD_VELOCITY_RESULT: return fb->velocity_bc;
        case GD_PRESSURE_RESULT: return PICT_BC_NEUMANN;
        case GD_SCALAR:
        case GD_SCALAR_RESULT: return fixed_scalar_bc(fb, ch);
        default: return PICT_BC_DIRICHLET;
    }
}

/* getBlockDataNeighbor :2518-2557 (zeroBound false; across a connection it reads the
   connected block at border offset 1, as the reference does) */
static real data_neighbor(const Ctx *c, int bi, I4 pos, int dir, int type) {
    const pict_block_desc *b = blk(c, bi);
    int dim = dir >> 1;
    I4 t = pos;
    int tb = bi;
    if (at_bound(pos, dir, b)) {
        const pict_boundary_desc *bd = &b->bounds[dir];
        switch (bd->type) {
            case PICT_BOUNDARY_FIXED: return fixed_data(c, b, t, dir, type);
            case PICT_BOUNDARY_CONNECTED: t = connected_pos(c, t, dim, bd, 1); tb = bd->connected_block; break;
            case PICT_BOUNDARY_PERIODIC: t.a[dim] = (dir & 1) ? 0 : b->size[dim] - 1; break;
            default: return 0;
        }
    } else {
        t.a[dim] += face_sign(dir);
    }
    return block_data(c, tb, t, type);
}
/* getBlockDataNeighborDiagonal :2614-2656: two steps, the prescribed direction last */
static real data_neighbor_diag(const Ctx *c, int bi, I4 pos, int dir1, int dir2, int type) {
    const int d1e = prescribed(dir1, blk(c, bi));
    const int dirs[2] = {d1e ? dir2 : dir1, d1e ? dir1 : dir2};
    int tb = bi;
    I4 t = pos;
    for (int i = 0; i < 2; ++i) {
        const pict_block_desc *b = blk(c, tb);
        int bound = dirs[i], dim = bound >> 1;
        if (at_bound(t, bound, b)) {
            const pict_boundary_desc *bd = &b->bounds[bound];
            switch (bd->type) {
                case PICT_BOUNDARY_FIXED: return fixed_data(c, b, t, bound, type);
                case PICT_BOUNDARY_CONNECTED: t = connected_pos(c, t, dim, bd, 1); tb = bd->connected_block; break;
                case PICT_BOUNDARY_PERIODIC: t.a[dim] = (bound & 1) ? 0 : b->size[dim] - 1; break;
                default: return 0;
            }
        } else {
            t.a[dim] += face_sign(bound);
        }
    }
    return block_data(c, tb, t, type);
}

/* getCornerValue :2742-2859: the mean of the cells around the corner of `pos` between faces dir1
   and dir2, walking around it in both directions (to depth maxDepth), through connections (border
   offset 1, as the reference) and periodic wraps; a FIXED boundary ends the walk with the boundary
   value (num = 0) and its condition type. */
typedef struct { real data; int num; int btype; } Corner;
typedef struct { int d1, d2, bi; I4 pos; } Walker;
static Corner corner_value(const Ctx *c, int bi, I4 pos, int dir1, int dir2, int inc0, int inc1, int max_depth,
                           int type) {
    real data = 0;
    int num = 0, btype = PICT_BC_DIRICHLET;
    if (inc0) data += block_data(c, bi, pos, type);
    ++num;
    Walker w[2] = {{dir1, dir2, bi, pos}, {dir2, dir1, bi, pos}};
    for (int depth = 1; depth <= max_depth; ++depth) {
        for (int d = 0; d < 2; ++d) {
            const pict_block_desc *wb = blk(c, w[d].bi);
            const int axis = w[d].d1 >> 1, fs = face_sign(w[d].d1);
            if (at_bound(w[d].pos, w[d].d1, wb)) {
                const pict_boundary_desc *bd = &wb->bounds[w[d].d1];
                if (bd->type == PICT_BOUNDARY_FIXED) {
                    const int at2 = at_bound(w[d].pos, w[d].d2, wb);
                    btype = fixed_type(wb, w[d].d1, type, w[d].pos.a[3]);
                    Corner r;
                    r.num = 0; r.btype = btype;
                    if (at2) {
                        r.data = fixed_data(c, wb, w[d].pos, w[d].d1, type);
                    } else {
                        I4 np = w[d].pos;
                        np.a[w[d].d2 >> 1] += face_sign(w[d].d2);
                        r.data = (fixed_data(c, wb, w[d].pos, w[d].d1, type) + fixed_data(c, wb, np, w[d].d1, type)) *
                                 (real)0.5;
                    }
                    return r;
                } else if (bd->type == PICT_BOUNDARY_CONNECTED) {
                    w[d].bi = bd->connected_block;
                    w[d].pos.a[axis] += fs;
                    w[d].pos = connected_pos(c, w[d].pos, axis, bd, 1);
                    const int o1 = w[d].d1;
                    w[d].d1 = connected_dir(c, w[d].d2, axis, bd);
                    w[d].d2 = connected_dir(c, o1, axis, bd) ^ 1;
                } else if (bd->type == PICT_BOUNDARY_PERIODIC) {
                    w[d].pos.a[axis] = (w[d].d1 & 1) ? 0 : wb->size[axis] - 1;
                    const int o1 = w[d].d1;
                    w[d].d1 = w[d].d2;
                    w[d].d2 = o1 ^ 1;
                }
            } else {
                w[d].pos.a[axis] += fs;
                const int o1 = w[d].d1;
                w[d].d1 = w[d].d2;
                w[d].d2 = o1 ^ 1;
            }
            if (w[d].bi == w[d ^ 1].bi && i4_eq(w[d].pos, w[d ^ 1].pos)) goto done;
            if (depth > 1 || inc1) data += block_data(c, w[d].bi, w[d].pos, type);
            ++num;
        }
    }
done:;
    Corner r;
    r.data = data / (real)num; r.num = num; r.btype = btype;
    return r;
}

/* interpolateNonOrthoLaplaceComponents :1911-1986 (cell metrics): out[face*(D-1) + i-1] is the
   face value of alpha_{axis, (axis+i)%D} (times 1/A and the viscosity when asked).  Faces of a
   cell at either end of an axis whose face on that axis is prescribed are left 0, as in the
   reference's condition (0 < x < size-1 || face not prescribed). */
/* useFaceTransform for the non-orthogonal pressure terms (:1962-1972, :3066-3068, :3128-3146): set only around the
   pressure stages that take the option, read by interp_nonortho and nonortho_rhs */
static int g_face_coefficients = 0;

/* getLaplaceCoefficientsSingleFace :1320-1343: det Minv[cc].Minv[axis] of the face's own transform (the unit vector
   along the face's axis without a transform) */
static real lap_single_face(const Ctx *c, const pict_block_desc *b, I4 pos, int face, int cc) {
    const int axis = face >> 1;
    if (!b->has_transform) return cc == axis ? (real)1 : (real)0;
    const real *T = face_T(c, b, pos, face);
    real s = 0;
    for (int j = 0; j < c->D; ++j) s += T_MINV(c, T, cc, j) * T_MINV(c, T, axis, j);
    return T_DET(c, T) * s;
}

static void interp_nonortho(const Ctx *c, int bi, I4 pos, int with_visc, int with_a, real *out) {
    const pict_block_desc *b = blk(c, bi);
    const real *A = (const real *)c->d->A;
    real aP[6];
    lap_full(c, b, pos, aP);
    const real one = 1;
    const real raP = (with_a ? one / A[global_index(pos, b)] : one) * (with_visc ? viscosity(c, b, pos) : one);
    for (int face = 0; face < 2 * c->D; ++face) {
        const int axis = face >> 1, up = face & 1;
        if ((0 < pos.a[axis] && pos.a[axis] < b->size[axis] - 1) || !prescribed(face, b)) {
            I4 pn = pos; pn.a[3] = axis;
            int nbi = bi;
            real aN[3];
            if (at_bound(pos, face, b) && b->bounds[face].type == PICT_BOUNDARY_CONNECTED) {
                nbi = b->bounds[face].connected_block;
                pn = connected_pos_ch(c, pn, axis, &b->bounds[face]);
                lap_single_axis(c, blk(c, nbi), pn, axis, aN);
            } else {
                if (!up && pos.a[axis] == 0 && b->bounds[face].type == PICT_BOUNDARY_PERIODIC) pn.a[axis] = b->size[axis] - 1;
                else if (up && pos.a[axis] == b->size[axis] - 1 && b->bounds[face].type == PICT_BOUNDARY_PERIODIC) pn.a[axis] = 0;
                else pn.a[axis] = pos.a[axis] + face_sign(face);
                lap_single_axis(c, b, pn, axis, aN);
            }
            pn.a[3] = 0;
            const pict_block_desc *nb = blk(c, nbi);
            const real raN = (with_a ? one / A[global_index(pn, nb)] : one) * (with_visc ? viscosity(c, nb, pn) : one);
            for (int i = 1; i < c->D; ++i) {
                const int oa = (axis + i) % c->D;
                out[face * (c->D - 1) + i - 1] = g_face_coefficients
                    ? lap_single_face(c, b, pos, face, oa) * (raP + raN) * (real)0.5
                    : (aP[lap_ci(axis, oa)] * raP + aN[oa] * raN) * (real)0.5;
            }
        } else {
            for (int i = 1; i < c->D; ++i) out[face * (c->D - 1) + i - 1] = 0;
        }
    }
}

/* Non-orthogonal matrix entries of one face of a row (PISO_build_matrix :3738-3800 with sign -1,
   PISO_build_pressure_matrix :4877-4932 with sign +1): the corners of `face` with its tangential
   faces, weighted by the interpolated cross coefficient and the corner's cell count.  `rv` holds
   the row in slot order (0 = diagonal, face+1 = neighbour); corner_type is the GridDataType of
   the corner walk; grad_always: every boundary corner is a gradient (Neumann) corner. */
static void nonortho_matrix_face(const Ctx *c, int bi, I4 pos, int face, const real *ai, int flags, real sign,
                                 int corner_type, int grad_always, real *rv) {
    const int D = c->D, dim = face >> 1, fs = face_sign(face);
    const int nbrs = flags & PICT_NON_ORTHO_DIRECT_MATRIX, ctr = flags & PICT_NON_ORTHO_CENTER_MATRIX;
    for (int i = 1; i < D; ++i) {
        const int ta = (dim + i) % D;
        const real alpha = ai[face * (D - 1) + i - 1];
        if (alpha == 0) continue;
        for (int tu = 0; tu < 2; ++tu) {
            const int tf = (ta << 1) | tu, tfs = face_sign(tf);
            const Corner cv = corner_value(c, bi, pos, face, tf, 0, 0, 2, corner_type);
            if (cv.num < 1) {
                if (grad_always || cv.btype == PICT_BC_NEUMANN) {
                    const real v = (real)(fs * tfs) * alpha * (real)0.25;
                    if (ctr) rv[0] += sign * 3 * v;
                    if (nbrs) { rv[face + 1] += sign * 3 * v; rv[(tf ^ 1) + 1] -= sign * v; }
                }
            } else {
                const real norm = (real)(1.0 / (double)cv.num);
                const real v = (real)(fs * tfs) * alpha * norm;
                if (ctr) rv[0] += sign * v;
                if (nbrs) { rv[face + 1] += sign * v; rv[tf + 1] += sign * v; }
            }
        }
    }
}

/* getNonOrthoLaplaceRHS_v2 :3033-3190 (no face transform): the explicit non-orthogonal part of the
   Laplacian of `type` at cell `pos` (w = component), from the previous iterate. */
static real nonortho_rhs(const Ctx *c, int bi, I4 pos, int flags, int type, int with_visc, int with_a) {
    if (!(flags & (PICT_NON_ORTHO_DIRECT_RHS | PICT_NON_ORTHO_DIAGONAL_RHS))) return 0;
    const pict_block_desc *b = blk(c, bi);
    const real *A = (const real *)c->d->A;
    const int D = c->D;
    real S = 0;
    real aP[6];
    lap_full(c, b, pos, aP);
    const real one = 1;
    const real raP = (with_a ? one / A[global_index(pos, b)] : one) * (with_visc ? viscosity(c, b, pos) : one);
    for (int face = 0; face < 2 * D; ++face) {
        const int axis = face >> 1, fs = face_sign(face);
        if (at_bound(pos, face, b) && prescribed(face, b)) {
            if (fixed_type(b, face, type, pos.a[3]) == PICT_BC_NEUMANN || gd_is_pressure(type)) continue;
            real bA[3];
            if (g_face_coefficients) for (int cc = 0; cc < D; ++cc) bA[cc] = lap_single_face(c, b, pos, face, cc);
            else lap_neighbour_boundary(c, &b->bounds[face], pos, axis, bA);
            for (int i = 1; i < D; ++i) {
                const int ta = (axis + i) % D;
                const int tlo = pos.a[ta] == 0, thi = pos.a[ta] == b->size[ta] - 1;
                I4 lo = pos, hi = pos;
                real df = (real)0.5;
                if (!tlo) lo.a[ta] -= 1;
                if (!thi) hi.a[ta] += 1;
                if (tlo || thi) df = 1;
                const real tdg = df * (fixed_data(c, b, hi, face, type) - fixed_data(c, b, lo, face, type));
                const real bv = with_visc ? viscosity(c, b, pos) : one;
                S -= (real)fs * bA[ta] * tdg * bv;
            }
            continue;
        }
        const NCell n = resolve_neighbor(c, bi, pos, face);
        real aN[3];
        lap_single_axis_neighbor(c, &n, face, aN);
        const pict_block_desc *nb = blk(c, n.bi);
        const real raN = (with_a ? one / A[global_index(n.pos, nb)] : one) * (with_visc ? viscosity(c, nb, n.pos) : one);
        for (int i = 1; i < D; ++i) {
            const int ta = (axis + i) % D;
            const real fa = g_face_coefficients ? lap_single_face(c, b, pos, face, ta) * (raP + raN) * (real)0.5
                                                : (aP[lap_ci(axis, ta)] * raP + aN[ta] * raN) * (real)0.5;
            real tdg = 0;
            for (int tu = 0; tu < 2; ++tu) {
                const int tf = (ta << 1) | tu, tfs = face_sign(tf);
                const Corner cv = corner_value(c, bi, pos, face, tf, 0, (flags & PICT_NON_ORTHO_DIRECT_RHS) != 0, 2, type);
                if (cv.num == 0 && (cv.btype == PICT_BC_NEUMANN || gd_is_pressure(type))) {
                    const int to = tf ^ 1, tofs = -tfs;
                    if (flags & PICT_NON_ORTHO_DIRECT_RHS) {
                        tdg += (real)tfs * data_neighbor(c, bi, pos, face, type) * (real)0.75;
                        tdg += (real)tofs * data_neighbor(c, bi, pos, to, type) * (real)0.25;
                    }
                    if (flags & PICT_NON_ORTHO_DIAGONAL_RHS)
                        tdg += (real)tofs * data_neighbor_diag(c, bi, pos, face, to, type) * (real)0.25;
                } else {
                    tdg += (real)tfs * cv.data;
                }
            }
            S -= (real)fs * fa * tdg;
        }
    }
    return S;
}

/* PISO_build_matrix :3601-3865: the velocity matrix (sc = 0), or the matrix of passive-scalar channel ch (sc = 1,
   forPassiveScalar: scalar viscosity :3683, :3726-3727, cross coefficients without viscosity times the scalar
   viscosity :3742-3755, the channel's boundary types for corners :3746-3750 and prescribed faces :3815-3826) */
static int advection_matrix(const pict_domain_desc *d, double time_step, int flags, int sc, int ch) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
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
            const real visc_p = sc ? scalar_viscosity(&c, ch) : viscosity(&c, b, pos);
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
                    real visc_n = sc ? scalar_viscosity(&c, ch) : viscosity(&c, blk(&c, nb), t);
                    idx[face + 1] = global_index(t, blk(&c, nb));
                    real vc = (alpha_p[dim] * visc_p + alpha * visc_n) * (real)0.5;
                    diag += vc; rv[face + 1] -= vc;
                    if (c.D > 1 && (flags & (PICT_NON_ORTHO_DIRECT_MATRIX | PICT_NON_ORTHO_CENTER_MATRIX))) {
                        real ai[12];
                        interp_nonortho(&c, bi, pos, !sc, 0, ai);
                        I4 cp = pos;
                        if (sc) {
                            for (int k = 0; k < 2 * c.D * (c.D - 1); ++k) ai[k] *= visc_p;
                            cp.a[3] = ch;
                        }
                        rv[0] = diag;
                        nonortho_matrix_face(&c, bi, cp, face, ai, flags, (real)-1, sc ? GD_SCALAR : GD_VELOCITY, 0, rv);
                        diag = rv[0];
                    }
                } else {
                    int bc = sc ? fixed_scalar_bc(&b->bounds[face], ch) : b->bounds[face].velocity_bc;
                    int slip = bc == PICT_BC_DIRICHLET ? 0 : 1;
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

int oracle_setup_advection_matrix(const pict_domain_desc *d, double time_step, int flags) {
    return advection_matrix(d, time_step, flags, 0, 0);
}
int oracle_setup_advection_matrix_scalar(const pict_domain_desc *d, double time_step, int flags, int channel) {
    if (d->scalar_channels < 1 || channel < 0 || channel >= d->scalar_channels) {
        snprintf(g_err, sizeof g_err, "passive scalar channel %d out of range", channel);
        return PICT_ERR_INVALID;
    }
    return advection_matrix(d, time_step, flags, 1, channel);
}

/* kPISO_build_scalar_advection_RHS :4080-4190: scalar_rhs of every channel */
int oracle_setup_advection_scalar(const pict_domain_desc *d, double time_step, int flags) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    const real dt = (real)time_step;
    real *rhs = (real *)d->scalar_rhs;
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        #pragma omp parallel for schedule(static)
        for (int f = 0; f < b->stride[3]; ++f) {
            I4 pos = unflat(f, b);
            const real det = cell_det(&c, b, pos);
            for (int ch = 0; ch < d->scalar_channels; ++ch) {
                I4 t = pos; t.a[3] = ch;
                const real visc = scalar_viscosity(&c, ch);
                real r = det * ((const real *)b->scalar)[flat4(t, b->stride)] / dt;
                for (int face = 0; face < 2 * c.D; ++face) {
                    if (!at_bound(pos, face, b) || b->bounds[face].type != PICT_BOUNDARY_FIXED) continue;
                    const pict_boundary_desc *fb = &b->bounds[face];
                    const int ax = face >> 1;
                    const real normal = (face & 1) ? 1 : -1;
                    I4 bp = pos; bp.a[ax] = 0;
                    const real sv = fixed_scalar(fb, bp, ch);    /* value (Dirichlet) or gradient (Neumann) */
                    bp.a[3] = ax;
                    const real alpha = laplace_orth_fixed(&c, fb, bp);
                    const real flux = fixed_contravariant(&c, fb, bp) * normal;
                    r -= sv * flux;
                    if (fixed_scalar_bc(fb, ch) == PICT_BC_DIRICHLET) r += sv * visc * 2 * alpha;
                    else r += sv * visc;
                }
                if (c.D > 1) r -= nonortho_rhs(&c, bi, t, flags, GD_SCALAR_RESULT, 0, 0) * visc;
                r /= det;
                rhs[(size_t)ch * c.N + b->global_offset + f] = r;
            }
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

/* kPISO_build_advection_RHS :4281-4385 */
int oracle_setup_advection_velocity(const pict_domain_desc *d, double time_step, int flags, int apply_pressure_gradient) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
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
                r -= nonortho_rhs(&c, bi, t, flags, GD_VELOCITY_RESULT, 1, 0);
                r /= det;
                r += velocity_source(b, t);
                if (apply_pressure_gradient) r -= pg[dd] * dt;
                rhs[(size_t)dd * c.N + b->global_offset + f] = r;
            }
        }
    }
    return PICT_OK;
}

/* PISO_build_pressure_matrix :4797-4963 (no face transform) */
/* use_face: the face-transform coefficient det|Minv[axis]|^2 of the face times the mean of the two 1/A (:4857-4865,
   orthogonal terms only) */
int oracle_setup_pressure_matrix(const pict_domain_desc *d, int flags, int use_face) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    if (use_face && c.D > 1 && (flags & (PICT_NON_ORTHO_DIRECT_MATRIX | PICT_NON_ORTHO_CENTER_MATRIX)))
        for (int bi = 0; bi < d->num_blocks; ++bi)
            if (d->blocks[bi].has_transform && !d->blocks[bi].face_transform) {
                snprintf(g_err, sizeof g_err, "block %d: face coefficients need face transforms", bi);
                return PICT_ERR_INVALID;
            }
    g_face_coefficients = use_face;
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
                if (use_face) coef = laplace_orth_face(&c, b, pos, face) * ((real)0.5 * (raP + raN));
                rv[0] -= coef; rv[face + 1] += coef; idx[face + 1] = gn;
                if (c.D > 1 && (flags & (PICT_NON_ORTHO_DIRECT_MATRIX | PICT_NON_ORTHO_CENTER_MATRIX))) {
                    real ai[12];
                    interp_nonortho(&c, bi, pos, 0, 1, ai);
                    nonortho_matrix_face(&c, bi, pos, face, ai, flags, (real)1, GD_IS_FIXED, 1, rv);
                }
            }
            int start = d->csr_row[g];
            int n = 0; for (int s = 0; s < 7; ++s) n += idx[s] >= 0;
            for (int i = 0; i < n; ++i) { int s = lowest(idx); Pv[start + i] = rv[s]; idx[s] = -1; }
        }
    }
    g_face_coefficients = 0;
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

/* slerp2Halfpoint :937-971 with eps 1e-5: the direction of the mean of the two vectors, each scaled by the first
   one's inverse length as the reference does (the direction is that of v1 + v2; the smallness test sees the scale),
   with the mean of their lengths; the plain mean when a vector or the direction is too short */
static void slerp2_half(const real *v1, const real *v2, int D, real *out) {
    const real eps = (real)1e-5;
    real s1 = 0, s2 = 0;
    for (int k = 0; k < D; ++k) { s1 += v1[k] * v1[k]; s2 += v2[k] * v2[k]; }
    if (s1 < eps || s2 < eps) { for (int k = 0; k < D; ++k) out[k] = (v1[k] + v2[k]) * (real)0.5; return; }
    const real r1 = (real)1 / sqrt(s1);
    real n[3], sn = 0;
    for (int k = 0; k < D; ++k) { n[k] = (v1[k] * r1 + v2[k] * r1) * (real)0.5; sn += n[k] * n[k]; }
    if (sn < eps) { for (int k = 0; k < D; ++k) out[k] = (v1[k] + v2[k]) * (real)0.5; return; }
    const real rn = (real)1 / sqrt(sn);
    const real m = (real)(((double)sqrt(s1) + (double)sqrt(s2)) * 0.5);
    for (int k = 0; k < D; ++k) out[k] = n[k] * rn * m;
}

/* computeFluxesWithFaceTransforms :1703-1785: per face, the neighbour's vector of vg (a fixed face: the boundary
   velocity) blended with the cell's by slerp2_half (the reference does so at fixed faces too), projected on the face's
   det * Minv[axis] */
static void fluxes_face(const Ctx *c, const pict_block_desc *b, I4 pos, real *fl, const real *vg) {
    const int D = c->D, g = global_index(pos, b);
    real vc[3];
    for (int k = 0; k < D; ++k) vc[k] = vg[g + (size_t)c->N * k];
    for (int face = 0; face < 2 * D; ++face) {
        const int dim = face >> 1, up = face & 1;
        real vn[3];
        I4 t = pos; t.a[3] = dim;
        if (at_bound(pos, face, b)) {
            const pict_boundary_desc *fb = &b->bounds[face];
            if (fb->type == PICT_BOUNDARY_FIXED) {
                I4 bp = pos; bp.a[dim] = 0; bp.a[3] = 0;
                for (int k = 0; k < D; ++k) vn[k] = fixed_vel(c, fb, bp, k);
            } else if (fb->type == PICT_BOUNDARY_CONNECTED) {
                const pict_block_desc *ob = blk(c, fb->connected_block);
                I4 o = connected_pos(c, t, dim, fb, 0);
                const int gn = global_index(o, ob);
                for (int k = 0; k < D; ++k) vn[k] = vg[gn + (size_t)c->N * k];
            } else {
                t.a[dim] = up ? 0 : b->size[dim] - 1;
                const int gn = global_index(t, b);
                for (int k = 0; k < D; ++k) vn[k] = vg[gn + (size_t)c->N * k];
            }
        } else {
            t.a[dim] += up ? 1 : -1;
            const int gn = global_index(t, b);
            for (int k = 0; k < D; ++k) vn[k] = vg[gn + (size_t)c->N * k];
        }
        real v[3];
        slerp2_half(vc, vn, D, v);
        const real *F = face_T(c, b, pos, face);
        real s = 0;
        for (int j = 0; j < D; ++j) s += T_MINV(c, F, dim, j) * v[j];
        fl[face] = T_DET(c, F) * s;
    }
}

/* k_computePressureRHSdivergenceFromFlux :5374-5419 (use_face: computeFluxesWithFaceTransforms; every block needs
   face transforms) */
int oracle_setup_pressure_rhs_div(const pict_domain_desc *d, double time_step, int time_step_norm, int use_face) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    const real dt = (real)time_step;
    if (use_face)
        for (int bi = 0; bi < d->num_blocks; ++bi)
            if (!d->blocks[bi].has_face_transform || !d->blocks[bi].face_transform) {
                snprintf(g_err, sizeof g_err, "block %d: fluxes with face transforms need face transforms", bi);
                return PICT_ERR_INVALID;
            }
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        #pragma omp parallel for schedule(static)
        for (int f = 0; f < b->stride[3]; ++f) {
            real fl[6];
            if (use_face) fluxes_face(&c, b, unflat(f, b), fl, (const real *)d->pressure_rhs);
            else fluxes_nd(&c, bi, unflat(f, b), fl, (const real *)d->pressure_rhs);
            real div = 0;
            for (int dd = 0; dd < c.D; ++dd) div += fl[2 * dd + 1] - fl[2 * dd];
            if (time_step_norm) div /= dt;
            ((real *)d->pressure_rhs_div)[f + b->global_offset] = div;
        }
    }
    return PICT_OK;
}

/* k_pressureRHSaddNonOrthoComponents :5455-5477: pressure_rhs_div += the explicit non-orthogonal
   part of the pressure Laplacian from the previous pressure result */
int oracle_pressure_rhs_add_non_ortho(const pict_domain_desc *d, int flags) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    if (!(flags & (PICT_NON_ORTHO_DIRECT_RHS | PICT_NON_ORTHO_DIAGONAL_RHS))) return PICT_OK;
    real *div = (real *)d->pressure_rhs_div;
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        #pragma omp parallel for schedule(static)
        for (int f = 0; f < b->stride[3]; ++f) {
            real S = div[f + b->global_offset];
            S += nonortho_rhs(&c, bi, unflat(f, b), flags, GD_PRESSURE_RESULT, 0, 1);
            div[f + b->global_offset] = S;
        }
    }
    return PICT_OK;
}

/* the same with the face coefficients of useFaceTransform (:3128-3146); every transformed block needs face
   transforms */
int oracle_pressure_rhs_add_non_ortho_face(const pict_domain_desc *d, int flags) {
    for (int bi = 0; bi < d->num_blocks; ++bi)
        if (d->blocks[bi].has_transform && !d->blocks[bi].face_transform) {
            snprintf(g_err, sizeof g_err, "block %d: face coefficients need face transforms", bi);
            return PICT_ERR_INVALID;
        }
    g_face_coefficients = 1;
    const int st = oracle_pressure_rhs_add_non_ortho(d, flags);
    g_face_coefficients = 0;
    return st;
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

/* getBlockDataGradient :2982-3025 for the velocity: per axis (value of the upper neighbour - value of the
 * lower neighbour) / distance in cells, the distance 2 less 0.5 for a Dirichlet face (boundary value half a
 * cell away) or 1 for a Neumann face (the cell's own value); a connected face gives the neighbour block's
 * cell with the same component index (computeConnectedPos :327-347 leaves w as it is); a periodic face wraps.
 * With a transform, out_j = sum_i g_i Minv_ij (matmul(vector, matrix), transformations.h:397-405). */
static void velocity_gradient(const Ctx *c, const pict_block_desc *b, I4 pos, int comp, real *out) {
    int D = c->D;
    real cg[3];
    pos.a[3] = comp;
    real own = ((const real *)b->velocity)[flat4(pos, b->stride)];
    for (int dim = 0; dim < D; ++dim) {
        real dist = 2.0, acc = 0.0;
        for (int up = 0; up < 2; ++up) {
            int face = 2 * dim + up;
            real v;
            if (at_bound(pos, face, b) && prescribed(face, b)) {
                const pict_boundary_desc *fb = &b->bounds[face];
                if (fb->velocity_bc == PICT_BC_DIRICHLET) {
                    I4 bp = pos;
                    bp.a[dim] = 0;
                    v = fixed_vel(c, fb, bp, comp);
                    dist -= 0.5;
                } else {
                    v = own;
                    dist -= 1.0;
                }
            } else {
                const pict_block_desc *nb = b;
                I4 t = pos;
                if (at_bound(pos, face, b) && b->bounds[face].type == PICT_BOUNDARY_CONNECTED) {
                    nb = blk(c, b->bounds[face].connected_block);
                    t = connected_pos(c, pos, dim, &b->bounds[face], 0);
                } else if (at_bound(pos, face, b)) {
                    t.a[dim] = up ? 0 : b->size[dim] - 1;
                } else {
                    t.a[dim] += up ? 1 : -1;
                }
                v = ((const real *)nb->velocity)[flat4(t, nb->stride)];
            }
            acc += up ? v : -v;
        }
        cg[dim] = acc / dist;
    }
    if (b->has_transform) {
        const real *T = cell_T(c, b, pos);
        for (int j = 0; j < D; ++j) {
            real s = 0;
            for (int i = 0; i < D; ++i) s += T_MINV(c, T, i, j) * cg[i];
            out[j] = s;
        }
    } else {
        for (int j = 0; j < D; ++j) out[j] = cg[j];
    }
}

/* k_computeSpatialVelocityGradients :6445-6467: out[(comp*D + i)*N + g] = d u_comp / d x_i */
int oracle_spatial_velocity_gradients(const pict_domain_desc *d, real *out) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        for (int f = 0; f < b->stride[3]; ++f) {
            I4 pos = unflat(f, b);
            for (int comp = 0; comp < c.D; ++comp) {
                real g[3];
                velocity_gradient(&c, b, pos, comp, g);
                for (int i = 0; i < c.D; ++i) out[(size_t)(comp * c.D + i) * c.N + b->global_offset + f] = g[i];
            }
        }
    }
    return PICT_OK;
}

/* k_SGSviscosityIncompressibleSmagorinsky :6898-6952: coefficient * Delta^2 * |S|, |S| = sqrt(2 sum S_ij^2) over
 * i <= j with the off-diagonal terms doubled (S = (G + G^T)/2, G_ij = d u_i / d x_j), Delta^2 the largest squared
 * column norm of M (1 without a transform) */
int oracle_sgs_smagorinsky(const pict_domain_desc *d, double coefficient, real *out) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    int D = c.D;
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        for (int f = 0; f < b->stride[3]; ++f) {
            I4 pos = unflat(f, b);
            real G[3][3];
            for (int comp = 0; comp < D; ++comp) velocity_gradient(&c, b, pos, comp, G[comp]);
            real s2 = 0;
            for (int i = 0; i < D; ++i)
                for (int j = i; j < D; ++j) {
                    real sv = (real)0.5 * (G[i][j] + G[j][i]);
                    sv *= sv;
                    if (i != j) sv *= 2;
                    s2 += sv;
                }
            real norm = sqrt(2 * s2);
            real delta = 1;
            if (b->has_transform) {
                const real *T = cell_T(&c, b, pos);
                delta = 0;
                for (int i = 0; i < D; ++i) {
                    real col = 0;
                    for (int r = 0; r < D; ++r) col += T[r * D + i] * T[r * D + i];
                    if (col > delta) delta = col;
                }
            }
            out[b->global_offset + f] = (real)coefficient * delta * norm;
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

/* getPressureGradientFVM :991-1124, modes 0 (the mean of the two cells' fluxes det * Minv[dim] * p) and 3 (face
   metrics times the mean pressure).  A fixed face extrapolates the pressure from the cell and its inner neighbour:
   half a cell with the boundary's transform (mode 0), a whole cell then averaged with the cell's (mode 3).  Across a
   connection the neighbour contributes its own axis' row, negated when both faces are on the same side. */
static void pressure_gradient_fvm(const Ctx *c, const pict_block_desc *b, I4 pos, int mode, real *g) {
    if (!b->has_transform) { pressure_gradient(c, b, pos, g); return; }
    const int D = c->D;
    pos.a[3] = 0;
    const real *TP = cell_T(c, b, pos);
    const real pP = ((const real *)b->pressure)[flat4(pos, b->stride)];
    real acc[3] = {0, 0, 0};
    for (int face = 0; face < 2 * D; ++face) {
        const int dim = face >> 1, up = face & 1;
        const real sgn = up ? 1 : -1;
        const real *T;
        int axn = dim;
        real tsign = 1, p;
        I4 t = pos;
        if (at_bound(pos, face, b)) {
            const pict_boundary_desc *fb = &b->bounds[face];
            if (fb->type == PICT_BOUNDARY_FIXED) {
                t.a[dim] = 0;
                T = bound_T(c, fb, t);
                t.a[dim] = pos.a[dim] - (up ? 1 : -1);
                const real pin = ((const real *)b->pressure)[flat4(t, b->stride)];
                if (mode != 3) {
                    p = pP + (pP - pin) * (real)0.5;
                    for (int j = 0; j < D; ++j) acc[j] += T_MINV(c, T, dim, j) * (T_DET(c, T) * p) * sgn;
                    continue;
                }
                p = pP + (pP - pin);
            } else if (fb->type == PICT_BOUNDARY_CONNECTED) {
                const pict_block_desc *ob = blk(c, fb->connected_block);
                t.a[3] = dim;
                I4 o = connected_pos_ch(c, t, dim, fb);
                axn = o.a[3];
                if ((fb->axes[0] & 1) == up) tsign = -1;
                o.a[3] = 0;
                T = cell_T(c, ob, o);
                p = ((const real *)ob->pressure)[flat4(o, ob->stride)];
            } else if (fb->type == PICT_BOUNDARY_PERIODIC) {
                t.a[dim] = up ? 0 : b->size[dim] - 1;
                T = cell_T(c, b, t);
                p = ((const real *)b->pressure)[flat4(t, b->stride)];
            } else {
                continue;
            }
        } else {
            t.a[dim] += up ? 1 : -1;
            T = cell_T(c, b, t);
            p = ((const real *)b->pressure)[flat4(t, b->stride)];
        }
        if (mode == 3) {
            const real *F = face_T(c, b, pos, face);
            for (int j = 0; j < D; ++j) acc[j] += T_MINV(c, F, dim, j) * (T_DET(c, F) * (p + pP) * (real)0.5 * sgn);
        } else {
            for (int j = 0; j < D; ++j)
                acc[j] += (T_MINV(c, TP, dim, j) * (T_DET(c, TP) * pP) + T_MINV(c, T, axn, j) * (T_DET(c, T) * p * tsign))
                          * ((real)0.5 * sgn);
        }
    }
    const real rdet = (real)1 / T_DET(c, TP);
    for (int j = 0; j < D; ++j) g[j] = acc[j] * rdet;
}

/* PISO_update_velocity_PressureFVM :6013-6049 (CorrectVelocity versions 5 and 6, :6180-6191) */
int oracle_correct_velocity_fvm(const pict_domain_desc *d, double time_step, int mode, int time_step_norm) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    const real dt = (real)time_step;
    const real *A = (const real *)d->A, *prhs = (const real *)d->pressure_rhs;
    real *vr = (real *)d->velocity_result;
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        if (!b->has_transform) continue;
        if (mode == 3 && !b->face_transform) {
            snprintf(g_err, sizeof g_err, "block %d: finite-volume gradient with face metrics needs face transforms", bi);
            return PICT_ERR_INVALID;
        }
        for (int f = 0; f < 2 * c.D; ++f)
            if (b->bounds[f].type == PICT_BOUNDARY_FIXED && !b->bounds[f].transform) {
                snprintf(g_err, sizeof g_err, "block %d face %d: fixed boundary without a transform", bi, f);
                return PICT_ERR_INVALID;
            }
    }
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        #pragma omp parallel for schedule(static)
        for (int f = 0; f < b->stride[3]; ++f) {
            I4 pos = unflat(f, b);
            int g = b->global_offset + f;
            real rDiag = 1 / A[g];
            real pg[3];
            pressure_gradient_fvm(&c, b, pos, mode, pg);
            for (int dd = 0; dd < c.D; ++dd) {
                real u = -rDiag * pg[dd];
                if (time_step_norm) u *= dt;
                vr[g + (size_t)c.N * dd] = u + prhs[g + (size_t)c.N * dd];
            }
        }
    }
    return PICT_OK;
}

/* getTransformMetricOrthogonal :1181-1192: Minv[w][w] of a cell (1 without a transform) */
static real metric_orth(const Ctx *c, const pict_block_desc *b, I4 p) {
    if (!b->has_transform) return 1;
    const int comp = p.a[3];
    const real *T = cell_T(c, b, p);
    return T_MINV(c, T, comp, comp);
}
/* PISO_update_velocity_v4_orthogonal :6052-6160 (CorrectVelocity version 4): per component a face flux from the
   two cells' covariant H/A and pressure difference, a quarter per side, over det*t.  The cell's own metric t is
   read at channel 0 for every component, as the reference's unflattened index carries w = 0; the neighbours' at
   the component's channel (remapped across connections).  Prescribed faces add nothing; the time step always
   scales the pressure term. */
int oracle_correct_velocity_v4(const pict_domain_desc *d, double time_step) {
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
            const real ra = 1 / A[g];
            const real p = ((const real *)b->pressure)[f];
            const real det = cell_det(&c, b, pos);
            const real t = metric_orth(&c, b, pos);
            const real cc = det * t;
            for (int dim = 0; dim < c.D; ++dim) {
                const size_t gc = g + (size_t)c.N * dim;
                const real hbyA = prhs[gc];
                real vel = 0;
                for (int up = 0; up < 2; ++up) {
                    const int face = 2 * dim + up;
                    if (at_bound(pos, face, b) && prescribed(face, b)) continue;
                    const pict_block_desc *nb = b;
                    I4 q = pos;
                    q.a[3] = dim;
                    if (at_bound(pos, face, b) && b->bounds[face].type == PICT_BOUNDARY_CONNECTED) {
                        nb = blk(&c, b->bounds[face].connected_block);
                        q = connected_pos_ch(&c, q, dim, &b->bounds[face]);
                    } else if (at_bound(pos, face, b)) {
                        q.a[dim] = up ? 0 : b->size[dim] - 1;
                    } else {
                        q.a[dim] += up ? 1 : -1;
                    }
                    const int gn = global_index(q, nb);
                    const real hbyAN = prhs[gn + (size_t)c.N * q.a[3]];
                    const real tN = metric_orth(&c, nb, q);
                    q.a[3] = 0;
                    const real detN = cell_det(&c, nb, q);
                    const real pN = ((const real *)nb->pressure)[flat4(q, nb->stride)];
                    const real raN = 1 / A[gn];
                    const real ccN = detN * tN;
                    const real dp = up ? pN - p : p - pN;
                    vel += (cc * hbyA + ccN * hbyAN - (cc * t * ra + ccN * tN * raN) * dp * dt) * (real)0.25;
                }
                vr[gc] = vel / (det * t);
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
/* CopyScalarResult{To,From}Blocks :6543-6610 (grad = 1: CopyScalarResultGrad{To,From}Blocks :6737-6800) */
int oracle_copy_scalar_result(const pict_domain_desc *d, int to_blocks, int grad) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        for (int ch = 0; ch < d->scalar_channels; ++ch) {
            real *bv = (real *)(grad ? b->scalar_grad : b->scalar) + (size_t)ch * b->stride[3];
            real *gv = (real *)(grad ? d->scalar_result_grad : d->scalar_result) + (size_t)ch * c.N + b->global_offset;
            if (to_blocks) memcpy(bv, gv, sizeof(real) * b->stride[3]);
            else memcpy(gv, bv, sizeof(real) * b->stride[3]);
        }
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

/* ------------------------------------------------------------------ */
/* non-orthogonal terms: vector-Jacobian products                       */
/* ------------------------------------------------------------------ */

/* scatterBlockData :2340-2347 (gradient fields of the global results) */
static void scatter_block_data(const Ctx *c, int bi, I4 p, int type, real g) {
    const pict_block_desc *b = blk(c, bi);
    real *dst = NULL;
    switch (GD_BASE(type)) {
        case GD_VELOCITY_RESULT: dst = (real *)c->d->velocity_result_grad; break;
        case GD_PRESSURE_RESULT: dst = (real *)c->d->pressure_result_grad; break;
        case GD_SCALAR_RESULT: dst = (real *)c->d->scalar_result_grad; break;
        default: return;
    }
    if (dst) dst[(size_t)p.a[3] * c->N + global_index(p, b)] += g;
}
/* scatterFixedBoundaryData :2448-2478 (FixedBoundary velocity and passive-scalar gradients) */
static void scatter_fixed_data(const Ctx *c, const pict_block_desc *b, I4 p, int face, int type, real g) {
    (void)c;
    const pict_boundary_desc *fb = &b->bounds[face];
    if (GD_BASE(type) == GD_IS_FIXED || fb->type != PICT_BOUNDARY_FIXED) return;
    if (gd_is_scalar(type)) {
        if (!fb->scalar_grad) return;
        real *sdst = (real *)fb->scalar_grad;
        if (fb->scalar_static) { sdst[p.a[3]] += g; return; }
        p.a[face >> 1] = 0;
        sdst[flat4(p, fb->stride)] += g;
        return;
    }
    if (GD_BASE(type) != GD_VELOCITY_RESULT && GD_BASE(type) != GD_VELOCITY) return;
    if (!fb->velocity_grad) return;
    real *dst = (real *)fb->velocity_grad;
    if (fb->velocity_static) { dst[p.a[3]] += g; return; }
    p.a[face >> 1] = 0;
    dst[flat4(p, fb->stride)] += g;
}
/* scatterBlockDataNeighbor :2560-2602 (boundary values get no gradient) */
static void scatter_data_neighbor(const Ctx *c, int bi, I4 pos, int dir, int type, real g) {
    const pict_block_desc *b = blk(c, bi);
    int dim = dir >> 1, tb = bi;
    I4 t = pos;
    if (at_bound(pos, dir, b)) {
        const pict_boundary_desc *bd = &b->bounds[dir];
        switch (bd->type) {
            case PICT_BOUNDARY_CONNECTED: t = connected_pos(c, t, dim, bd, 1); tb = bd->connected_block; break;
            case PICT_BOUNDARY_PERIODIC: t.a[dim] = (dir & 1) ? 0 : b->size[dim] - 1; break;
            default: return;
        }
    } else {
        t.a[dim] += face_sign(dir);
    }
    scatter_block_data(c, tb, t, type, g);
}
/* scatterBlockDataNeighborDiagonal :2659-2700 */
static void scatter_data_neighbor_diag(const Ctx *c, int bi, I4 pos, int dir1, int dir2, int type, real g) {
    const int d1e = prescribed(dir1, blk(c, bi));
    const int dirs[2] = {d1e ? dir2 : dir1, d1e ? dir1 : dir2};
    int tb = bi;
    I4 t = pos;
    for (int i = 0; i < 2; ++i) {
        const pict_block_desc *b = blk(c, tb);
        int bound = dirs[i], dim = bound >> 1;
        if (at_bound(t, bound, b)) {
            const pict_boundary_desc *bd = &b->bounds[bound];
            switch (bd->type) {
                case PICT_BOUNDARY_CONNECTED: t = connected_pos(c, t, dim, bd, 1); tb = bd->connected_block; break;
                case PICT_BOUNDARY_PERIODIC: t.a[dim] = (bound & 1) ? 0 : b->size[dim] - 1; break;
                default: return;
            }
        } else {
            t.a[dim] += face_sign(bound);
        }
    }
    scatter_block_data(c, tb, t, type, g);
}
/* scatterCornerValue_GRAD :2869-2966: the gradient g of a corner value (with the forward walk's
   cell count) to the cells it averaged, or to the boundary values it took. */
static void scatter_corner_value(const Ctx *c, int bi, I4 pos, int dir1, int dir2, int inc0, int inc1, int max_depth,
                                 int type, real g, int num) {
    const int from_cells = num > 0;
    real dg = g;
    if (from_cells) dg /= (real)num;
    if (inc0) scatter_block_data(c, bi, pos, type, dg);
    Walker w[2] = {{dir1, dir2, bi, pos}, {dir2, dir1, bi, pos}};
    for (int depth = 1; depth <= max_depth; ++depth) {
        for (int d = 0; d < 2; ++d) {
            const pict_block_desc *wb = blk(c, w[d].bi);
            const int axis = w[d].d1 >> 1, fs = face_sign(w[d].d1);
            if (at_bound(w[d].pos, w[d].d1, wb)) {
                const pict_boundary_desc *bd = &wb->bounds[w[d].d1];
                if (bd->type == PICT_BOUNDARY_FIXED) {
                    if (at_bound(w[d].pos, w[d].d2, wb)) {
                        scatter_fixed_data(c, wb, w[d].pos, w[d].d1, type, dg);
                    } else {
                        I4 np = w[d].pos;
                        np.a[w[d].d2 >> 1] += face_sign(w[d].d2);
                        dg *= (real)0.5;
                        scatter_fixed_data(c, wb, w[d].pos, w[d].d1, type, dg);
                        scatter_fixed_data(c, wb, np, w[d].d1, type, dg);
                    }
                    return;
                } else if (bd->type == PICT_BOUNDARY_CONNECTED) {
                    w[d].bi = bd->connected_block;
                    w[d].pos.a[axis] += fs;
                    w[d].pos = connected_pos(c, w[d].pos, axis, bd, 1);
                    const int o1 = w[d].d1;
                    w[d].d1 = connected_dir(c, w[d].d2, axis, bd);
                    w[d].d2 = connected_dir(c, o1, axis, bd) ^ 1;
                } else if (bd->type == PICT_BOUNDARY_PERIODIC) {
                    w[d].pos.a[axis] = (w[d].d1 & 1) ? 0 : wb->size[axis] - 1;
                    const int o1 = w[d].d1;
                    w[d].d1 = w[d].d2;
                    w[d].d2 = o1 ^ 1;
                }
            } else {
                w[d].pos.a[axis] += fs;
                const int o1 = w[d].d1;
                w[d].d1 = w[d].d2;
                w[d].d2 = o1 ^ 1;
            }
            if (w[d].bi == w[d ^ 1].bi && i4_eq(w[d].pos, w[d ^ 1].pos)) return;
            if (from_cells && (depth > 1 || inc1)) scatter_block_data(c, w[d].bi, w[d].pos, type, dg);
        }
    }
}

/* scatterNonOrthoLaplaceComponents_GRAD :1989-2098: gradients of the interpolated cross coefficients
   (ai_grad, laid out as interp_nonortho's output) to the viscosities and A of the cell and its
   face neighbours. */
static void scatter_interp_nonortho(const Ctx *c, int bi, I4 pos, const real *ai_grad, int with_visc, int with_a) {
    const pict_block_desc *b = blk(c, bi);
    const real *A = (const real *)c->d->A;
    real *Ag = (real *)c->d->A_grad;
    const int D = c->D;
    real aP[6];
    lap_full(c, b, pos, aP);
    const real one = 1;
    real raP_grad = 0;
    for (int face = 0; face < 2 * D; ++face) {
        const int axis = face >> 1, up = face & 1;
        if (!((0 < pos.a[axis] && pos.a[axis] < b->size[axis] - 1) || !prescribed(face, b))) continue;
        I4 pn = pos; pn.a[3] = axis;
        int nbi = bi;
        real aN[3];
        if (at_bound(pos, face, b) && b->bounds[face].type == PICT_BOUNDARY_CONNECTED) {
            nbi = b->bounds[face].connected_block;
            pn = connected_pos_ch(c, pn, axis, &b->bounds[face]);
            lap_single_axis(c, blk(c, nbi), pn, axis, aN);
        } else {
            if (!up && pos.a[axis] == 0 && b->bounds[face].type == PICT_BOUNDARY_PERIODIC) pn.a[axis] = b->size[axis] - 1;
            else if (up && pos.a[axis] == b->size[axis] - 1 && b->bounds[face].type == PICT_BOUNDARY_PERIODIC) pn.a[axis] = 0;
            else pn.a[axis] = pos.a[axis] + face_sign(face);
            lap_single_axis(c, b, pn, axis, aN);
        }
        pn.a[3] = 0;
        const pict_block_desc *nb = blk(c, nbi);
        real raN_grad = 0;
        for (int i = 1; i < D; ++i) {
            const int oa = (axis + i) % D;
            const real g = ai_grad[face * (D - 1) + i - 1];
            raP_grad += aP[lap_ci(axis, oa)] * (real)0.5 * g;
            raN_grad += aN[oa] * (real)0.5 * g;
        }
        const int gn = global_index(pn, nb);
        const real raN = with_a ? one / A[gn] : one;
        if (with_visc) {
            real t = raN_grad;
            if (with_a) t *= raN;
            scatter_viscosity(c, nb, pn, t);
        }
        if (with_a) {
            if (with_visc) raN_grad *= viscosity(c, nb, pn);
            if (Ag) Ag[gn] += -raN_grad * raN * raN;
        }
    }
    const int g = global_index(pos, b);
    const real raP = with_a ? one / A[g] : one;
    if (with_visc) {
        real t = raP_grad;
        if (with_a) t *= raP;
        scatter_viscosity(c, b, pos, t);
    }
    if (with_a) {
        if (with_visc) raP_grad *= viscosity(c, b, pos);
        if (Ag) Ag[g] += -raP_grad * raP * raP;
    }
}

/* VJP of nonortho_matrix_face (PISO_build_matrix_GRAD :3948-4019, PISO_build_pressure_matrix_GRAD
   :5038-5104): accumulates into ai_grad the gradient of the face's cross coefficients from the row
   gradient cg (slot order, cg[0] the diagonal's). */
static void nonortho_matrix_face_grad(const Ctx *c, int bi, I4 pos, int face, const real *ai, int flags, real sign,
                                      int corner_type, int grad_always, const real *cg, real *ai_grad) {
    const int D = c->D, dim = face >> 1, fs = face_sign(face);
    const int nbrs = flags & PICT_NON_ORTHO_DIRECT_MATRIX, ctr = flags & PICT_NON_ORTHO_CENTER_MATRIX;
    for (int i = 1; i < D; ++i) {
        const int ta = (dim + i) % D;
        if (ai[face * (D - 1) + i - 1] == 0) continue;
        real ag = 0;
        for (int tu = 0; tu < 2; ++tu) {
            const int tf = (ta << 1) | tu, tfs = face_sign(tf);
            const Corner cv = corner_value(c, bi, pos, face, tf, 0, 0, 2, corner_type);
            if (cv.num < 1) {
                if (grad_always || cv.btype == PICT_BC_NEUMANN) {
                    real vg = 0;
                    if (ctr) vg += 3 * cg[0];
                    if (nbrs) { vg += 3 * cg[face + 1]; vg -= cg[(tf ^ 1) + 1]; }
                    ag += (real)(fs * tfs) * (real)0.25 * (sign * vg);
                }
            } else {
                const real norm = (real)(1.0 / (double)cv.num);
                real vg = 0;
                if (ctr) vg += cg[0];
                if (nbrs) { vg += cg[face + 1]; vg += cg[tf + 1]; }
                ag += (real)(fs * tfs) * norm * (sign * vg);
            }
        }
        ai_grad[face * (D - 1) + i - 1] += ag;
    }
}

/* scatterNonOrthoLaplaceRHS_v2_GRAD :3210-3380 (face coefficients under g_face_coefficients, :3294-3312,
   :3364-3368): the VJP of nonortho_rhs for the
   output gradient s_grad; `type` is the forward GridDataType (gradients go to the matching
   result-gradient vector, FixedBoundary velocity gradients, viscosities and A). */
static void nonortho_rhs_grad(const Ctx *c, int bi, I4 pos, int flags, int type, int with_visc, int with_a,
                              real s_grad) {
    if (!(flags & (PICT_NON_ORTHO_DIRECT_RHS | PICT_NON_ORTHO_DIAGONAL_RHS))) return;
    const pict_block_desc *b = blk(c, bi);
    const real *A = (const real *)c->d->A;
    real *Ag = (real *)c->d->A_grad;
    const int D = c->D;
    const int direct = (flags & PICT_NON_ORTHO_DIRECT_RHS) != 0;
    real aP[6];
    lap_full(c, b, pos, aP);
    const real one = 1;
    const int g = global_index(pos, b);
    const real raP = with_a ? one / A[g] : one;
    const real viscP = with_visc ? viscosity(c, b, pos) : one;
    const real rvP = raP * viscP;
    const int diffable = with_visc || with_a;
    real raP_grad = 0;
    for (int face = 0; face < 2 * D; ++face) {
        const int axis = face >> 1, fs = face_sign(face);
        if (at_bound(pos, face, b) && prescribed(face, b)) {
            if (fixed_type(b, face, type, pos.a[3]) == PICT_BC_NEUMANN || gd_is_pressure(type)) continue;
            real bA[3];
            if (g_face_coefficients) for (int cc = 0; cc < D; ++cc) bA[cc] = lap_single_face(c, b, pos, face, cc);
            else lap_neighbour_boundary(c, &b->bounds[face], pos, axis, bA);
            for (int i = 1; i < D; ++i) {
                const int ta = (axis + i) % D;
                const int tlo = pos.a[ta] == 0, thi = pos.a[ta] == b->size[ta] - 1;
                I4 lo = pos, hi = pos;
                real df = (real)0.5;
                if (!tlo) lo.a[ta] -= 1;
                if (!thi) hi.a[ta] += 1;
                if (tlo || thi) df = 1;
                const real tgg = -(real)fs * bA[ta] * s_grad * df;
                if (with_visc) {
                    const real tdg = fixed_data(c, b, hi, face, type) - fixed_data(c, b, lo, face, type);
                    scatter_viscosity(c, b, pos, tgg * tdg);
                }
                const real bv = with_visc ? viscosity(c, b, pos) : one;
                scatter_fixed_data(c, b, hi, face, type, tgg * bv);
                scatter_fixed_data(c, b, lo, face, type, -tgg * bv);
            }
            continue;
        }
        const NCell n = resolve_neighbor(c, bi, pos, face);
        real aN[3];
        lap_single_axis_neighbor(c, &n, face, aN);
        const pict_block_desc *nb = blk(c, n.bi);
        const int gn = global_index(n.pos, nb);
        const real raN = with_a ? one / A[gn] : one;
        const real viscN = with_visc ? viscosity(c, nb, n.pos) : one;
        const real rvN = raN * viscN;
        real raN_grad = 0;
        for (int i = 1; i < D; ++i) {
            const int ta = (axis + i) % D;
            const real cP = g_face_coefficients ? lap_single_face(c, b, pos, face, ta) : aP[lap_ci(axis, ta)];
            const real cN = g_face_coefficients ? cP : aN[ta];
            const real fa = (cP * rvP + cN * rvN) * (real)0.5;
            const real tgg = -s_grad * (real)fs * fa;
            real tdg = 0;
            for (int tu = 0; tu < 2; ++tu) {
                const int tf = (ta << 1) | tu, tfs = face_sign(tf);
                const Corner cv = corner_value(c, bi, pos, face, tf, 0, direct, 2, type);
                if (cv.num == 0 && (cv.btype == PICT_BC_NEUMANN || gd_is_pressure(type))) {
                    const int to = tf ^ 1, tofs = -tfs;
                    if (direct) {
                        if (diffable) tdg += (real)tfs * data_neighbor(c, bi, pos, face, type) * (real)0.75;
                        scatter_data_neighbor(c, bi, pos, face, type, tgg * (real)tfs * (real)0.75);
                        if (diffable) tdg += (real)tofs * data_neighbor(c, bi, pos, to, type) * (real)0.25;
                        scatter_data_neighbor(c, bi, pos, to, type, tgg * (real)tofs * (real)0.25);
                    }
                    if (flags & PICT_NON_ORTHO_DIAGONAL_RHS) {
                        if (diffable) tdg += (real)tofs * data_neighbor_diag(c, bi, pos, face, to, type) * (real)0.25;
                        scatter_data_neighbor_diag(c, bi, pos, face, to, type, tgg * (real)tofs * (real)0.25);
                    }
                } else {
                    if (diffable) tdg += (real)tfs * cv.data;
                    scatter_corner_value(c, bi, pos, face, tf, 0, direct, 2, type, tgg * (real)tfs, cv.num);
                }
            }
            if (diffable) {
                const real fag = -s_grad * (real)fs * tdg;
                raP_grad += cP * (real)0.5 * fag;
                raN_grad += cN * (real)0.5 * fag;
            }
        }
        if (with_visc) scatter_viscosity(c, nb, n.pos, raN_grad * raN);
        if (with_a && Ag) Ag[gn] += raN_grad * viscN * (-raN * raN);
    }
    if (with_visc) scatter_viscosity(c, b, pos, raP_grad * raP);
    if (with_a && Ag) Ag[g] += raP_grad * viscP * (-raP * raP);
}

/* PISO_build_matrix_GRAD :3869-4075 */
/* scatterViscosity_GRAD(forPassiveScalar) :1834-1849: into the scalar viscosity gradient of channel ch, or the
   domain viscosity gradient when there is none */
static void scatter_scalar_viscosity(const Ctx *c, int ch, real g) {
    real *dst = (real *)c->d->scalar_viscosity_grad;
    if (dst) dst += c->d->scalar_viscosity_static ? 0 : ch;
    else dst = (real *)c->d->viscosity_grad;
    if (dst) *dst += g;
}

/* PISO_build_matrix_GRAD :3869-4075: the velocity matrix (sc = 0) or that of passive-scalar channel ch (sc = 1:
   viscosity gradients to the scalar viscosity, the cross coefficients' gradient times their viscosity-free value
   :4004-4011, the channel's boundary types) */
static int advection_matrix_grad(const pict_domain_desc *d, double time_step, int flags, int sc, int ch) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
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
            const int no = c.D > 1 && (flags & (PICT_NON_ORTHO_DIRECT_MATRIX | PICT_NON_ORTHO_CENTER_MATRIX));
            real ai[12], ai_grad[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            if (no) interp_nonortho(&c, bi, pos, 0, 0, ai);
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
                    if (sc) scatter_scalar_viscosity(&c, ch, (real)0.5 * alpha * vcg);
                    else scatter_viscosity(&c, blk(&c, nb), t, (real)0.5 * alpha * vcg);
                    if (no) {
                        real cgd[7];
                        for (int k = 0; k < 7; ++k) cgd[k] = cg[k];
                        cgd[0] = diag_grad;
                        I4 cp = pos;
                        if (sc) cp.a[3] = ch;
                        nonortho_matrix_face_grad(&c, bi, cp, face, ai, flags, (real)-1, sc ? GD_SCALAR : GD_VELOCITY, 0,
                                                  cgd, ai_grad);
                    }
                } else {
                    int bc = sc ? fixed_scalar_bc(&b->bounds[face], ch) : b->bounds[face].velocity_bc;
                    real slip = bc == PICT_BC_DIRICHLET ? 0 : 1;
                    visc_grad += (1 - slip) * 2 * alpha_p[dim] * diag_grad;
                }
            }
            if (no) {
                if (sc) {
                    for (int k = 0; k < 2 * c.D * (c.D - 1); ++k) visc_grad += ai_grad[k] * ai[k];
                } else {
                    scatter_interp_nonortho(&c, bi, pos, ai_grad, 1, 0);
                }
            }
            if (b->velocity_grad) scatter_fluxes(&c, bi, pos, fg, NULL);
            if (sc) scatter_scalar_viscosity(&c, ch, visc_grad);
            else scatter_viscosity(&c, b, pos, visc_grad);
        }
    }
    return PICT_OK;
}
int oracle_setup_advection_matrix_grad(const pict_domain_desc *d, double time_step, int flags) {
    return advection_matrix_grad(d, time_step, flags, 0, 0);
}
int oracle_setup_advection_matrix_scalar_grad(const pict_domain_desc *d, double time_step, int flags, int channel) {
    if (d->scalar_channels < 1 || channel < 0 || channel >= d->scalar_channels) {
        snprintf(g_err, sizeof g_err, "passive scalar channel %d out of range", channel);
        return PICT_ERR_INVALID;
    }
    return advection_matrix_grad(d, time_step, flags, 1, channel);
}

/* kPISO_build_scalar_advection_RHS_GRAD :4193-4278 */
int oracle_setup_advection_scalar_grad(const pict_domain_desc *d, double time_step, int flags) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    const real dt = (real)time_step;
    const real *rg = (const real *)d->scalar_rhs_grad;
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        for (int f = 0; f < b->stride[3]; ++f) {
            I4 pos = unflat(f, b);
            real det = cell_det(&c, b, pos);
            for (int ch = 0; ch < d->scalar_channels; ++ch) {
                I4 t = pos; t.a[3] = ch;
                const real visc = scalar_viscosity(&c, ch);
                real G = rg[(size_t)ch * c.N + b->global_offset + f] / det;
                real vg = 0;
                if (c.D > 1 && (flags & (PICT_NON_ORTHO_DIRECT_RHS | PICT_NON_ORTHO_DIAGONAL_RHS))) {
                    nonortho_rhs_grad(&c, bi, t, flags, GD_SCALAR_RESULT, 0, 0, -G * visc);
                    vg -= G * nonortho_rhs(&c, bi, t, flags, GD_SCALAR_RESULT, 0, 0);
                }
                for (int face = 0; face < 2 * c.D; ++face) {
                    if (!at_bound(pos, face, b) || b->bounds[face].type != PICT_BOUNDARY_FIXED) continue;
                    const pict_boundary_desc *fb = &b->bounds[face];
                    const int ax = face >> 1;
                    const real normal = (face & 1) ? 1 : -1;
                    I4 bp = pos; bp.a[ax] = 0;
                    const real sv = fixed_scalar(fb, bp, ch);
                    bp.a[3] = ax;
                    const real alpha = laplace_orth_fixed(&c, fb, bp);
                    const real flux = fixed_contravariant(&c, fb, bp) * normal;
                    real s_grad = -flux * G;
                    if (fixed_scalar_bc(fb, ch) == PICT_BC_DIRICHLET) {
                        s_grad += visc * 2 * alpha * G;
                        vg += sv * 2 * alpha * G;
                    } else {
                        s_grad += visc * G;
                        vg += sv * G;
                    }
                    scatter_fixed_contravariant(&c, fb, bp, -sv * G * normal);
                    scatter_fixed_data(&c, b, t, face, GD_SCALAR, s_grad);
                }
                scatter_scalar_viscosity(&c, ch, vg);
                if (b->scalar_grad) ((real *)b->scalar_grad)[flat4(t, b->stride)] = G * det / dt;
            }
        }
    }
    return PICT_OK;
}

/* kPISO_build_advection_RHS_GRAD :4388-4476 */
int oracle_setup_advection_velocity_grad(const pict_domain_desc *d, double time_step, int flags) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
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
                nonortho_rhs_grad(&c, bi, t, flags, GD_VELOCITY_RESULT, 1, 0, -G);
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
    const real *Pg = (const real *)d->P_grad, *A = (const real *)d->A;
    real *Ag = (real *)d->A_grad;
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        for (int f = 0; f < b->stride[3]; ++f) {
            I4 pos = unflat(f, b);
            int g = b->global_offset + f;
            int at[7]; slot_positions(&c, bi, pos, g, at);
            real raP_grad = 0;
            const int no = c.D > 1 && (flags & (PICT_NON_ORTHO_DIRECT_MATRIX | PICT_NON_ORTHO_CENTER_MATRIX));
            real ai[12], ai_grad[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}, pcg[7];
            if (no) {
                interp_nonortho(&c, bi, pos, 0, 0, ai);
                for (int k = 0; k < 7; ++k) pcg[k] = at[k] >= 0 ? Pg[at[k]] : 0;
            }
            for (int face = 0; face < 2 * c.D; ++face) {
                if (at_bound(pos, face, b) && prescribed(face, b)) continue;
                if (no) nonortho_matrix_face_grad(&c, bi, pos, face, ai, flags, (real)1, GD_IS_FIXED, 1, pcg, ai_grad);
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
            if (no) scatter_interp_nonortho(&c, bi, pos, ai_grad, 0, 1);
        }
    }
    return PICT_OK;
}

/* k_pressureRHSaddNonOrthoComponents_GRAD :5479-5494: pressure_rhs_div_grad -> pressure_result_grad,
   A_grad (+=) */
int oracle_pressure_rhs_add_non_ortho_grad(const pict_domain_desc *d, int flags) {
    Ctx c; int st = init_ctx(&c, d); if (st) return st;
    if (!(flags & (PICT_NON_ORTHO_DIRECT_RHS | PICT_NON_ORTHO_DIAGONAL_RHS))) return PICT_OK;
    const real *dg = (const real *)d->pressure_rhs_div_grad;
    for (int bi = 0; bi < d->num_blocks; ++bi) {
        const pict_block_desc *b = blk(&c, bi);
        for (int f = 0; f < b->stride[3]; ++f)
            nonortho_rhs_grad(&c, bi, unflat(f, b), flags, GD_PRESSURE_RESULT, 0, 1, dg[f + b->global_offset]);
    }
    return PICT_OK;
}

/* the same with the face coefficients of useFaceTransform, which this VJP uses (:5491) */
int oracle_pressure_rhs_add_non_ortho_grad_face(const pict_domain_desc *d, int flags) {
    for (int bi = 0; bi < d->num_blocks; ++bi)
        if (d->blocks[bi].has_transform && !d->blocks[bi].face_transform) {
            snprintf(g_err, sizeof g_err, "block %d: face coefficients need face transforms", bi);
            return PICT_ERR_INVALID;
        }
    g_face_coefficients = 1;
    const int st = oracle_pressure_rhs_add_non_ortho_grad(d, flags);
    g_face_coefficients = 0;
    return st;
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

/* The matrix the preconditioned solve works on: A, or A^T built explicitly (rows of A^T with ascending columns). */
typedef struct { int n; int *row, *col; real *val; } Csr;

static Csr csr_effective(const pict_csr_desc *A, int transpose) {
    Csr B; int n = A->n, nnz = A->row[n];
    const real *av = (const real *)A->value;
    B.n = n;
    B.row = (int *)calloc((size_t)n + 1, sizeof(int));
    B.col = (int *)malloc(sizeof(int) * (size_t)(nnz ? nnz : 1));
    B.val = (real *)malloc(sizeof(real) * (size_t)(nnz ? nnz : 1));
    if (!transpose) {
        memcpy(B.row, A->row, sizeof(int) * ((size_t)n + 1));
        memcpy(B.col, A->index, sizeof(int) * (size_t)nnz);
        memcpy(B.val, av, sizeof(real) * (size_t)nnz);
        return B;
    }
    for (int k = 0; k < nnz; ++k) B.row[A->index[k] + 1]++;
    for (int i = 0; i < n; ++i) B.row[i + 1] += B.row[i];
    int *fill = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    memcpy(fill, B.row, sizeof(int) * ((size_t)n + 1));
    for (int i = 0; i < n; ++i)  /* rows of A in order: each row of A^T gets its columns ascending */
        for (int k = A->row[i]; k < A->row[i + 1]; ++k) {
            int d = fill[A->index[k]]++;
            B.col[d] = i;
            B.val[d] = av[k];
        }
    free(fill);
    return B;
}

/* ILU(0) (csrilu02, bicgstab_solver_kernel.cu:201-209 with the factors' roles of :73-77): unit-lower L and upper U
   on the matrix's own pattern with (LU)_ij = a_ij on that pattern, row by row (IKJ: each earlier row k of the
   lower part scales a_ik by 1/u_kk and subtracts l_ik * u_kj from the row's later entries that row k has too).
   lu holds L below and U on and above the diagonal; diag[i] is the position of (i,i).  Returns the first row
   without a diagonal entry, or -1. */
static int ilu0(const Csr *B, real *lu, int *diag) {
    for (int i = 0; i < B->n; ++i) {
        diag[i] = -1;
        for (int k = B->row[i]; k < B->row[i + 1]; ++k) if (B->col[k] == i) diag[i] = k;
        if (diag[i] < 0) return i;
    }
    memcpy(lu, B->val, sizeof(real) * (size_t)B->row[B->n]);
    for (int i = 0; i < B->n; ++i)
        for (int kk = B->row[i]; kk < diag[i]; ++kk) {
            const int k = B->col[kk];
            lu[kk] /= lu[diag[k]];
            int q = diag[k] + 1;
            for (int jj = kk + 1; jj < B->row[i + 1]; ++jj) {
                const int j = B->col[jj];
                while (q < B->row[k + 1] && B->col[q] < j) ++q;
                if (q < B->row[k + 1] && B->col[q] == j) lu[jj] -= lu[kk] * lu[q];
            }
        }
    return -1;
}

/* out = (LU)^-1 in: forward substitution with the unit-lower L into tmp, back substitution with U. */
static void ilu_apply(const Csr *B, const real *lu, const int *diag, const real *in, real *tmp, real *out) {
    for (int i = 0; i < B->n; ++i) {
        real s = in[i];
        for (int k = B->row[i]; k < diag[i]; ++k) s -= lu[k] * tmp[B->col[k]];
        tmp[i] = s;
    }
    for (int i = B->n - 1; i >= 0; --i) {
        real s = tmp[i];
        for (int k = diag[i] + 1; k < B->row[i + 1]; ++k) s -= lu[k] * out[B->col[k]];
        out[i] = s / lu[diag[i]];
    }
}

/* bicgstabSolveGPU with preconditioner, bicgstab_solver_kernel.cu:222-365: right-preconditioned BiCGStab, the
   preconditioner applied to the search direction (pw = M^-1 p) and to the intermediate residual (s = M^-1 r); the
   stopping rule and the reported iterations as without preconditioner. */
static void bicg_ilu_one(const Csr *B, const real *lu, const int *diag, const real *f, real *x,
                         const pict_solve_opts *o, pict_solve_info *info, real *w) {
    const int n = B->n;
    real *r = w, *rw = w + n, *p = w + 2 * n, *pw = w + 3 * n, *v = w + 4 * n, *s = w + 5 * n, *t = w + 6 * n,
         *tmp = w + 7 * n;
    const Mat m = {B->row, B->col, B->val, n, 0};
    info->final_residual = 0; info->used_iterations = -1; info->converged = 0; info->is_finite_residual = 1;
    double rho = 0, rhop, alpha = 0, omega = 0, beta;
    spmv(&m, x, r);
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
        ilu_apply(B, lu, diag, p, tmp, pw);
        spmv(&m, pw, v);
        alpha = rho / dotd(rw, v, n);
        for (int i = 0; i < n; ++i) { r[i] -= (real)alpha * v[i]; x[i] += (real)alpha * pw[i]; }
        double nr = criterion(r, n, o->convergence_criterion);
        if (!isfinite(nr)) { info->converged = 0; info->used_iterations = it; info->final_residual = nr; info->is_finite_residual = 0; break; }
        info->used_iterations = it; info->final_residual = nr;
        if (nr < o->tolerance) { info->converged = 1; break; }
        ilu_apply(B, lu, diag, r, tmp, s);
        spmv(&m, s, t);
        omega = dotd(t, r, n) / dotd(t, t, n);
        for (int i = 0; i < n; ++i) { x[i] += (real)omega * s[i]; r[i] -= (real)omega * t[i]; }
        nr = criterion(r, n, o->convergence_criterion);
        info->final_residual = nr;
        if (nr < o->tolerance) { info->converged = 1; info->used_iterations = it + 1; break; }
    }
}

/* The ILU(0) factors of A (or of A^T with transpose_a) as one array on A's effective pattern: for the checks of
   the factorization itself. */
int oracle_ilu0(const pict_csr_desc *A, int transpose, real *lu_out) {
    if (A->dtype != ORACLE_DTYPE) { snprintf(g_err, sizeof g_err, "oracle: csr dtype mismatch"); return PICT_ERR_INVALID; }
    Csr B = csr_effective(A, transpose);
    int *diag = (int *)malloc(sizeof(int) * (size_t)A->n);
    int miss = ilu0(&B, lu_out, diag);
    free(diag); free(B.row); free(B.col); free(B.val);
    if (miss >= 0) { snprintf(g_err, sizeof g_err, "ILU(0): A(%d,%d) is missing", miss, miss); return PICT_ERR_INVALID; }
    return PICT_OK;
}

int oracle_solve_linear(const pict_csr_desc *A, const real *rhs, real *x, int nrhs, const pict_solve_opts *o, pict_solve_info *infos) {
    if (A->dtype != ORACLE_DTYPE) { snprintf(g_err, sizeof g_err, "oracle: csr dtype mismatch"); return PICT_ERR_INVALID; }
    if (o->matrix_rank_deficient) {
        snprintf(g_err, sizeof g_err, "oracle: option not restated"); return PICT_ERR_UNSUPPORTED;
    }
    if (o->use_bicg && o->bicg_with_preconditioner) {
        const int n = A->n;
        Csr B = csr_effective(A, o->transpose_a);
        real *lu = (real *)malloc(sizeof(real) * (size_t)(B.row[n] ? B.row[n] : 1));
        int *diag = (int *)malloc(sizeof(int) * (size_t)n);
        real *w = (real *)malloc(sizeof(real) * (size_t)n * 8);
        int miss = ilu0(&B, lu, diag);
        if (miss < 0)
            for (int k = 0; k < nrhs; ++k)
                bicg_ilu_one(&B, lu, diag, rhs + (size_t)k * n, x + (size_t)k * n, o, infos + k, w);
        free(w); free(diag); free(lu); free(B.row); free(B.col); free(B.val);
        if (miss >= 0) { snprintf(g_err, sizeof g_err, "ILU(0): A(%d,%d) is missing", miss, miss); return PICT_ERR_INVALID; }
        return PICT_OK;
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
