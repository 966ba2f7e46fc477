// pict_solve.hip -- CG and BiCGStab for the PISO linear systems on gfx950.
//
// One cooperative (persistent) launch per solve.  All right-hand sides of a
// call are iterated together, each with its own scalars and stopping test, so
// every right-hand side follows exactly the iteration it would follow alone
// (the reference solves them one after the other, cg_solver_kernel.cu:208-433,
// bicgstab_solver_kernel.cu:219-365).  Dot products, norms and the Krylov
// scalars are accumulated in FP64 whatever the vector precision.
//
// Grid-wide synchronisation is a "reduce barrier": every workgroup writes its
// FP64 partial sums, arrives on a device-scope counter, and the last workgroup
// to arrive reduces all partials in a fixed order (deterministic), updates the
// per-RHS state and releases the others.  The textbook iterations are
// re-associated so that a CG iteration needs 2 and a BiCGStab iteration 3 such
// barriers: the search direction update p = r + beta*p (BiCGStab: with the
// omega*v term) is folded into the following SpMV by evaluating it on the fly
// for the neighbouring entries, with double-buffered p (and v), using an
// explicit fma so that the on-the-fly value is bitwise the owner's value.
//
// Every wait is bounded: a barrier that spins longer than kSpinLimit sets an
// error flag, after which all workgroups leave the kernel.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "pict_host.hpp"
#include "pict_solver.hpp"

namespace pict {

constexpr int kSolveThreads = 256;
constexpr int kSlots = 8;              // FP64 partial slots per right-hand side
constexpr unsigned kSpinLimit = 1u << 23;
constexpr int kMaxRhs = 4;

// Control block, one 128-byte line per hot word (device memory, zeroed before launch).
struct Ctl {
    unsigned count;              // top-level arrivals (one per group)
    unsigned pad0[31];
    unsigned gen;                // barrier generation
    unsigned pad1[31];
    int err;                     // a wait exceeded its spin limit
    int zero_all;                // every right-hand side is zero
    unsigned pad2[30];
    unsigned group_count[8][32]; // per-group arrivals
};
constexpr int kGroups = 8;       // arrival tree fan-in (one group per XCD when blocks are dealt round-robin)

struct RhsState {
    double rho, rho_prev, alpha, omega, beta;
    double best_crit, last_crit;
    int active, converged, finite, used;
    int best_it, rising, save_best, restore_best;
    double final_res;
    double pad;
};

template <typename R>
struct SolveArgs {
    const int32_t *row, *col, *perm;  // perm != nullptr: use A^T (value of (j,i) for entry (i,j))
    const R *val;
    const R *b;
    R *x;
    R *r, *rhat, *p0, *p1, *v0, *v1, *s, *t, *best;  // each [nrhs*n]
    double *partials;                              // [nwg * nrhs * kSlots]
    Ctl *ctl;
    RhsState *st;                                  // [nrhs]
    int n, nrhs, maxit, conv, reset_steps, return_best, zero_shortcut;
    double tol;
    unsigned nwg;
};

// --------------------------------------------------------------------------
// workgroup reductions (wave64 shuffles + LDS), results broadcast to all lanes
// --------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}
// Reduce K values per thread over the workgroup; out[k] valid in all threads.
template <int K>
__device__ __forceinline__ void block_reduce(double (&v)[K], const bool (&is_max)[K]) {
    __shared__ double lds[kSolveThreads / 64][K];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double w = is_max[k] ? wave_max(v[k]) : wave_sum(v[k]);
        if (lane == 0) lds[wave][k] = w;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double a = lds[0][k];
#pragma unroll
        for (int w = 1; w < kSolveThreads / 64; ++w) a = is_max[k] ? fmax(a, lds[w][k]) : a + lds[w][k];
        v[k] = a;
    }
    __syncthreads();
}

// --------------------------------------------------------------------------
// grid reduce-barrier
// --------------------------------------------------------------------------
// Two-level arrival tree: workgroups of one group (blockIdx % G) count on their
// group's line; the last of each group arrives at the top counter; the last top
// arrival is the master.  Release fences publish this workgroup's writes before
// it arrives; the master and the waiters take one acquire fence after the wait
// (never inside the spin, which only does relaxed loads).
__device__ __forceinline__ bool arrive(Ctl *ctl, unsigned nwg, int &s_err) {
    __shared__ int s_last;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned G = nwg < kGroups ? nwg : kGroups;
        const unsigned g = blockIdx.x % G;
        const unsigned gsize = (nwg - g + G - 1) / G;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned prev = __hip_atomic_fetch_add(&ctl->group_count[g][0], 1u, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
        int last = 0;
        if (prev == gsize - 1) {
            __hip_atomic_store(&ctl->group_count[g][0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
            const unsigned top = __hip_atomic_fetch_add(&ctl->count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (top == G - 1) {
                __hip_atomic_store(&ctl->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                last = 1;
            }
        }
        s_last = last;
        s_err = 0;
    }
    __syncthreads();
    return s_last != 0;
}
// master: publish the new state and open the next generation
__device__ __forceinline__ void release(Ctl *ctl) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(&ctl->gen, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
}
__device__ __forceinline__ void wait_gen(Ctl *ctl, unsigned my_gen, int &s_err) {
    if (threadIdx.x == 0) {
        unsigned spins = 0;
        int err = 0;
        while (__hip_atomic_load(&ctl->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == my_gen) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > kSpinLimit) {
                __hip_atomic_store(&ctl->err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                err = 1;
                break;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (!err) err = __hip_atomic_load(&ctl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_err = err;
    }
    __syncthreads();
}

// slot layout of the partial sums (per workgroup and right-hand side)
enum { S_RR = 0, S_ASUM = 1, S_AMAX = 2, S_D0 = 3, S_D1 = 4, S_BMAX = 5 };

// Flush NS accumulators per RHS of this thread: workgroup-reduce and store
// this workgroup's partials for the slots `ids`.
template <typename R, int NR, int NS>
__device__ __forceinline__ void flush(const SolveArgs<R> &a, double (&acc)[NR][NS], const int (&ids)[NS],
                                      const bool (&ismax)[NS]) {
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        block_reduce<NS>(acc[k], ismax);
        if (threadIdx.x == 0)
#pragma unroll
            for (int s = 0; s < NS; ++s) a.partials[((size_t)blockIdx.x * a.nrhs + k) * kSlots + ids[s]] = acc[k][s];
    }
}

// Master step: reduce every workgroup's partials (fixed order) for the slots
// `ids`; red[k*kSlots + id] receives the totals (LDS, valid in all threads).
template <int NS>
__device__ __forceinline__ void master_reduce(const double *partials, unsigned nwg, int nrhs, const int (&ids)[NS],
                                              const bool (&ismax)[NS], double *red) {
    for (int k = 0; k < nrhs; ++k) {
        double v[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            double acc = 0.0;
            for (unsigned w = threadIdx.x; w < nwg; w += kSolveThreads) {
                const double pv = partials[((size_t)w * nrhs + k) * kSlots + ids[s]];
                acc = ismax[s] ? fmax(acc, pv) : acc + pv;
            }
            v[s] = acc;
        }
        block_reduce<NS>(v, ismax);
        if (threadIdx.x == 0)
#pragma unroll
            for (int s = 0; s < NS; ++s) red[k * kSlots + ids[s]] = v[s];
    }
    __syncthreads();
}

// criterion from (sum r^2, sum |r|, max |r|), ComputeConvergenceCriterion cg_solver_kernel.cu:77-110
__device__ __forceinline__ double criterion(int conv, double rr, double asum, double amax, int n) {
    switch (conv) {
        case PICT_CONV_NORM2: return sqrt(rr);
        case PICT_CONV_NORM2_NORMALIZED: return sqrt(rr) / sqrt((double)n);
        case PICT_CONV_ABS_SUM: return asum;
        case PICT_CONV_ABS_MEAN: return asum / n;
        default: return amax;
    }
}

template <typename R>
__device__ __forceinline__ R mat_val(const SolveArgs<R> &a, int k) {
    return a.perm ? a.val[a.perm[k]] : a.val[k];
}

template <int NR, int NS>
__device__ __forceinline__ void zero_acc(double (&acc)[NR][NS]) {
#pragma unroll
    for (int k = 0; k < NR; ++k)
#pragma unroll
        for (int s = 0; s < NS; ++s) acc[k][s] = 0.0;
}
// residual accumulators r^2, |r|, max|r| in slots 0..2 of acc[k]
template <typename R, int NR, int NS>
__device__ __forceinline__ void acc_res(double (&acc)[NR][NS], int k, R rv) {
    const double d = (double)rv;
    acc[k][0] += d * d;
    acc[k][1] += fabs(d);
    acc[k][2] = fmax(acc[k][2], fabs(d));
}

// Run one reduce-barrier: flush, arrive, master (reduce + `update`), release/wait.
#define PICT_REDUCE_BARRIER(NS, ACC, IDS, ISMAX, UPDATE)                                  \
    do {                                                                                  \
        flush<R, NRHS, NS>(a, ACC, IDS, ISMAX);                                           \
        if (arrive(a.ctl, a.nwg, s_err)) {                                                \
            master_reduce<NS>(a.partials, a.nwg, a.nrhs, IDS, ISMAX, red);                \
            if (threadIdx.x == 0) { UPDATE; }                                             \
            release(a.ctl);                                                               \
        } else {                                                                          \
            wait_gen(a.ctl, gen, s_err);                                                  \
        }                                                                                 \
        ++gen;                                                                            \
        if (s_err) return;                                                                \
    } while (0)

// --------------------------------------------------------------------------
// Flat reduce-barrier with replicated reduction (single-reduction CG and
// BiCGStab): every workgroup stores its block-reduced FP64 partials (double
// buffered by generation parity), publishes them with a release fence and one
// relaxed atomic increment; the last arrival opens the next generation.
// After the wait every workgroup reduces all partials itself in a fixed order,
// so all workgroups hold bitwise identical totals and run the same scalar
// update on their own LDS copy of the solver state.
// --------------------------------------------------------------------------
constexpr int kPartStride = kMaxRhs * kSlots;

// Wave64 reductions in DPP (quad and row rotations within 16-lane rows, then the
// four row totals in a fixed order): every lane of the wave gets the same total,
// and the result depends only on the inputs, so every workgroup reducing the same
// values gets bitwise the same result.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_d(double v, int lane) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                            __builtin_amdgcn_readlane(__double2loint(v), lane));
}
template <bool MAX>
__device__ __forceinline__ double wave_reduce_dpp(double v) {
    auto op = [](double x, double y) { return MAX ? fmax(x, y) : x + y; };
    v = op(v, dpp_d<0xB1>(v));   // quad_perm [1,0,3,2]
    v = op(v, dpp_d<0x4E>(v));   // quad_perm [2,3,0,1]
    v = op(v, dpp_d<0x124>(v));  // row_ror:4
    v = op(v, dpp_d<0x128>(v));  // row_ror:8
    return op(op(lane_d(v, 0), lane_d(v, 16)), op(lane_d(v, 32), lane_d(v, 48)));
}

template <int NR, int NS>
__device__ __forceinline__ void flat_reduce(double *part, Ctl *ctl, unsigned nwg, unsigned gen,
                                            double (&acc)[NR][NS], const int (&ids)[NS],
                                            const bool (&ismax)[NS], double *red, int &s_err) {
    __shared__ double wsum[kSolveThreads / 64][NR * NS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // 1. wave totals, then the workgroup total (thread 0, fixed order)
#pragma unroll
    for (int k = 0; k < NR; ++k)
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const double w = ismax[s] ? wave_reduce_dpp<true>(acc[k][s]) : wave_reduce_dpp<false>(acc[k][s]);
            if (lane == 0) wsum[wave][k * NS + s] = w;
        }
    __syncthreads();
    if (wave != 0) return;
    double *base = part + (size_t)(gen & 1) * nwg * kPartStride;
    // 2. publish this workgroup's partials and arrive
    if (lane == 0) {
        double *mine = base + (size_t)blockIdx.x * kPartStride;
#pragma unroll
        for (int k = 0; k < NR; ++k)
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                double t = wsum[0][k * NS + s];
#pragma unroll
                for (int w = 1; w < kSolveThreads / 64; ++w)
                    t = ismax[s] ? fmax(t, wsum[w][k * NS + s]) : t + wsum[w][k * NS + s];
                mine[k * kSlots + ids[s]] = t;
            }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned prev = __hip_atomic_fetch_add(&ctl->count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (prev == nwg - 1) {
            __hip_atomic_store(&ctl->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
            __hip_atomic_fetch_add(&ctl->gen, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    // 3. the whole wave waits for the generation to move (relaxed loads; one acquire after)
    unsigned spins = 0;
    int err = 0;
    while (__hip_atomic_load(&ctl->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
        __builtin_amdgcn_s_sleep(1);
        if (++spins > kSpinLimit) {
            if (lane == 0) __hip_atomic_store(&ctl->err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            err = 1;
            break;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    // 4. every workgroup's partials, reduced in the same fixed order by every workgroup
    double v[NR][NS];
#pragma unroll
    for (int k = 0; k < NR; ++k)
#pragma unroll
        for (int s = 0; s < NS; ++s) v[k][s] = 0.0;
    for (unsigned w = lane; w < nwg; w += 64) {
        const double *pw = base + (size_t)w * kPartStride;
#pragma unroll
        for (int k = 0; k < NR; ++k)
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const double pv = pw[k * kSlots + ids[s]];
                v[k][s] = ismax[s] ? fmax(v[k][s], pv) : v[k][s] + pv;
            }
    }
#pragma unroll
    for (int k = 0; k < NR; ++k)
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const double t = ismax[s] ? wave_reduce_dpp<true>(v[k][s]) : wave_reduce_dpp<false>(v[k][s]);
            if (lane == 0) red[k * kSlots + ids[s]] = t;
        }
    if (lane == 0) s_err = err;
}

// Reduce-barrier: flat_reduce leaves the totals and the error flag in LDS (written by
// thread 0, which then runs the scalar update); one workgroup barrier publishes both.
#define PICT_FLAT_BARRIER(NS, ACC, IDS, ISMAX, UPDATE)                                            \
    do {                                                                                          \
        flat_reduce<NRHS, NS>(a.partials, a.ctl, a.nwg, gen, ACC, IDS, ISMAX, red, s_err);       \
        ++gen;                                                                                    \
        if (threadIdx.x == 0 && !s_err) { UPDATE; }                                               \
        __syncthreads();                                                                          \
        if (s_err) return;                                                                        \
    } while (0)

// Matrix rows of one thread kept in registers across iterations: RPT rows,
// row q = i0 + q * stride (the grid-stride assignment), at most 7 entries.
constexpr int kMaxNz = 7;
constexpr int kRowsCached = 2;
template <typename R, int RPT>
struct RowCache {
    int len[RPT];
    int col[RPT][kMaxNz];
    R val[RPT][kMaxNz];
};

// ==========================================================================
// CG  (cgSolveGPU, cg_solver_kernel.cu:114-456)
// ==========================================================================
template <typename R, int NRHS>
__global__ __launch_bounds__(kSolveThreads) void k_cg(SolveArgs<R> a) {
    __shared__ double red[NRHS * kSlots];
    __shared__ int s_err;
    const int n = a.n;
    const size_t stride = (size_t)a.nwg * kSolveThreads;
    const size_t i0 = (size_t)blockIdx.x * kSolveThreads + threadIdx.x;
    unsigned gen = 0;
    constexpr int IDS_R[4] = {S_RR, S_ASUM, S_AMAX, S_BMAX};
    constexpr bool MAX_R[4] = {false, false, true, true};
    constexpr int IDS_RES[3] = {S_RR, S_ASUM, S_AMAX};
    constexpr bool MAX_RES[3] = {false, false, true};
    constexpr int IDS_D[1] = {S_D0};
    constexpr bool MAX_D[1] = {false};

    // ---- phase 0: r = b - A x, p_old = 0; rho = r.r ----
    {
        double acc[NRHS][4];
        zero_acc(acc);
        for (size_t i = i0; i < (size_t)n; i += stride) {
            const int rs = a.row[i], re = a.row[i + 1];
#pragma unroll
            for (int k = 0; k < NRHS; ++k) {
                
                const size_t o = (size_t)k * n;
                R s = 0;
                for (int e = rs; e < re; ++e) s += mat_val(a, e) * a.x[o + a.col[e]];
                const R bv = a.b[o + i];
                const R rv = bv - s;
                a.r[o + i] = rv;
                a.p0[o + i] = 0;
                acc_res(acc, k, rv);
                acc[k][3] = fmax(acc[k][3], fabs((double)bv));
            }
        }
        PICT_REDUCE_BARRIER(4, acc, IDS_R, MAX_R, {
            bool all_zero = true;
            for (int k = 0; k < NRHS; ++k) all_zero = all_zero && red[k * kSlots + S_BMAX] == 0.0;
            const int zero_all = (a.zero_shortcut && all_zero) ? 1 : 0;
            a.ctl->zero_all = zero_all;
            for (int k = 0; k < NRHS; ++k) {
                RhsState &st = a.st[k];
                st.rho = red[k * kSlots + S_RR];
                st.beta = 0;
                st.active = zero_all ? 0 : 1;
                st.converged = zero_all;
                st.finite = 1;
                st.used = -1;
                st.final_res = 0;
                st.best_crit = 0;
                st.last_crit = 0;
                st.best_it = -1;
                st.rising = 0;
                st.save_best = 0;
                st.restore_best = 0;
            }
        });
    }
    if (a.ctl->zero_all) {
        for (size_t i = i0; i < (size_t)n; i += stride)
            for (int k = 0; k < NRHS; ++k) a.x[(size_t)k * n + i] = 0;
        return;
    }

    for (int it = 0; it < a.maxit; ++it) {
        // ---- optional residual reset r = b - A x, p = r (cg_solver_kernel.cu:266-287) ----
        if (a.reset_steps > 0 && (it + 1) % a.reset_steps == 0) {
            bool act[NRHS];
#pragma unroll
            for (int k = 0; k < NRHS; ++k) act[k] = k < NRHS && a.st[k].active;
            double acc[NRHS][3];
            zero_acc(acc);
            for (size_t i = i0; i < (size_t)n; i += stride) {
                const int rs = a.row[i], re = a.row[i + 1];
#pragma unroll
                for (int k = 0; k < NRHS; ++k) {
                    if (!act[k]) continue;
                    const size_t o = (size_t)k * n;
                    R s = 0;
                    for (int e = rs; e < re; ++e) s += mat_val(a, e) * a.x[o + a.col[e]];
                    const R rv = a.b[o + i] - s;
                    a.r[o + i] = rv;
                    acc_res(acc, k, rv);
                }
            }
            PICT_REDUCE_BARRIER(3, acc, IDS_RES, MAX_RES, {
                for (int k = 0; k < NRHS; ++k)
                    if (a.st[k].active) {
                        a.st[k].rho = red[k * kSlots + S_RR];
                        a.st[k].beta = 0;
                    }
            });
        }
        R *pold = (it & 1) ? a.p1 : a.p0;
        R *pnew = (it & 1) ? a.p0 : a.p1;
        // ---- phase 1: p = r + beta p (evaluated on the fly), Ap = A p; p.Ap ----
        {
            bool act[NRHS], save[NRHS];
            R beta[NRHS];
#pragma unroll
            for (int k = 0; k < NRHS; ++k) {
                act[k] = k < NRHS && a.st[k].active;
                save[k] = act[k] && a.st[k].save_best;
                beta[k] = act[k] ? (R)a.st[k].beta : R(0);
            }
            double acc[NRHS][1];
            zero_acc(acc);
            for (size_t i = i0; i < (size_t)n; i += stride) {
                const int rs = a.row[i], re = a.row[i + 1];
#pragma unroll
                for (int k = 0; k < NRHS; ++k) {
                    if (!act[k]) continue;
                    const size_t o = (size_t)k * n;
                    if (save[k]) a.best[o + i] = a.x[o + i];
                    const R pn = fma(beta[k], pold[o + i], a.r[o + i]);
                    pnew[o + i] = pn;
                    R s = 0;
                    for (int e = rs; e < re; ++e) {
                        const size_t c = o + a.col[e];
                        s += mat_val(a, e) * fma(beta[k], pold[c], a.r[c]);
                    }
                    a.v0[o + i] = s;  // Ap
                    acc[k][0] += (double)pn * (double)s;
                }
            }
            PICT_REDUCE_BARRIER(1, acc, IDS_D, MAX_D, {
                for (int k = 0; k < NRHS; ++k)
                    if (a.st[k].active) {
                        a.st[k].alpha = a.st[k].rho / red[k * kSlots + S_D0];
                        a.st[k].save_best = 0;
                    }
            });
        }
        // ---- phase 2: x += alpha p, r -= alpha Ap; r.r, convergence ----
        {
            bool act[NRHS];
            R alpha[NRHS];
#pragma unroll
            for (int k = 0; k < NRHS; ++k) {
                act[k] = k < NRHS && a.st[k].active;
                alpha[k] = act[k] ? (R)a.st[k].alpha : R(0);
            }
            double acc[NRHS][3];
            zero_acc(acc);
            for (size_t i = i0; i < (size_t)n; i += stride) {
#pragma unroll
                for (int k = 0; k < NRHS; ++k) {
                    if (!act[k]) continue;
                    const size_t o = (size_t)k * n;
                    a.x[o + i] += alpha[k] * pnew[o + i];
                    const R rv = a.r[o + i] - alpha[k] * a.v0[o + i];
                    a.r[o + i] = rv;
                    acc_res(acc, k, rv);
                }
            }
            PICT_REDUCE_BARRIER(3, acc, IDS_RES, MAX_RES, {
                for (int k = 0; k < NRHS; ++k) {
                    RhsState &st = a.st[k];
                    if (!st.active) continue;
                    const double rr = red[k * kSlots + S_RR];
                    const double crit = criterion(a.conv, rr, red[k * kSlots + S_ASUM], red[k * kSlots + S_AMAX], n);
                    if (!isfinite(crit)) {
                        st.converged = 0; st.used = it; st.final_res = crit; st.finite = 0; st.active = 0;
                        continue;
                    }
                    if (a.return_best) {
                        if (it == 0 || crit < st.best_crit) { st.best_crit = crit; st.best_it = it; st.save_best = 1; }
                        if (it > 0 && crit >= st.last_crit) ++st.rising; else st.rising = 0;
                        st.last_crit = crit;
                    }
                    st.used = it;
                    st.final_res = crit;
                    if (crit < a.tol) { st.converged = 1; st.finite = 1; st.active = 0; continue; }
                    if (a.return_best && (it == a.maxit - 1 || st.rising >= 100)) {
                        st.restore_best = st.best_it != it;
                        st.converged = 0; st.used = st.best_it; st.final_res = st.best_crit; st.finite = 1;
                        st.active = 0;
                        continue;
                    }
                    st.rho_prev = st.rho;
                    st.rho = rr;
                    st.beta = st.rho / st.rho_prev;
                }
            });
        }
        bool any = false;
        for (int k = 0; k < NRHS; ++k) any = any || a.st[k].active;
        if (!any) break;
    }
    // restore best iterates (cg_solver_kernel.cu:399-412)
    for (int k = 0; k < NRHS; ++k) {
        if (!a.st[k].restore_best) continue;
        const size_t o = (size_t)k * n;
        for (size_t i = i0; i < (size_t)n; i += stride) a.x[o + i] = a.best[o + i];
    }
}

// ==========================================================================
// CG, single-reduction form (Chronopoulos & Gear 1989): the same Krylov
// iterates as cgSolveGPU (cg_solver_kernel.cu:114-456) in exact arithmetic,
// with (r,r) and (Ar,r) reduced together, so an iteration costs one
// reduce-barrier.  Per iteration i, for every own row m:
//   p = r + beta p;  s = w + beta s;  x += alpha p;  r' = r - alpha s;
//   w' = A r'  (r'_j of the neighbours evaluated on the fly from r, w, s of the
//   previous iteration, bitwise equal to the owner's value through explicit fma);
//   gamma' = (r', r'), delta' = (w', r');
//   beta' = gamma'/gamma, alpha' = gamma' / (delta' - beta' gamma' / alpha).
// r, w, s are double-buffered (neighbours read the previous values); p and x
// are only touched by their owner.
// ==========================================================================
template <typename R, int NRHS>
__global__ __launch_bounds__(kSolveThreads) void k_cg1(SolveArgs<R> a) {
    __shared__ double red[NRHS * kSlots];
    __shared__ RhsState sst[NRHS];
    __shared__ int s_err;
    const int n = a.n;
    const size_t stride = (size_t)a.nwg * kSolveThreads;
    const size_t i0 = (size_t)blockIdx.x * kSolveThreads + threadIdx.x;
    unsigned gen = 0;
    R *rb[2] = {a.r, a.rhat}, *wb[2] = {a.v0, a.v1}, *sb[2] = {a.s, a.t};
    R *p = a.p0;
    constexpr int IDS_R[4] = {S_RR, S_ASUM, S_AMAX, S_BMAX};
    constexpr bool MAX_R[4] = {false, false, true, true};
    constexpr int IDS_G[5] = {S_RR, S_ASUM, S_AMAX, S_D0, S_D1};
    constexpr bool MAX_G[5] = {false, false, true, false, false};
    int cur = 0;  // buffer holding r_i, w_i, s_{i-1}
    int zero_all = 0;

    // ---- phase I0: r0 = b - A x0; p = s = 0 ----
    {
        double acc[NRHS][4];
        zero_acc(acc);
        for (size_t i = i0; i < (size_t)n; i += stride) {
            const int rs = a.row[i], re = a.row[i + 1];
#pragma unroll
            for (int k = 0; k < NRHS; ++k) {
                
                const size_t o = (size_t)k * n;
                R s = 0;
                for (int e = rs; e < re; ++e) s += mat_val(a, e) * a.x[o + a.col[e]];
                const R bv = a.b[o + i];
                const R rv = bv - s;
                rb[0][o + i] = rv;
                p[o + i] = 0;
                sb[0][o + i] = 0;
                acc_res(acc, k, rv);
                acc[k][3] = fmax(acc[k][3], fabs((double)bv));
            }
        }
        PICT_FLAT_BARRIER(4, acc, IDS_R, MAX_R, {
            bool all_zero = true;
            for (int k = 0; k < NRHS; ++k) all_zero = all_zero && red[k * kSlots + S_BMAX] == 0.0;
            const int za = (a.zero_shortcut && all_zero) ? 1 : 0;
            for (int k = 0; k < NRHS; ++k) {
                RhsState &st = sst[k];
                st.rho = red[k * kSlots + S_RR];
                st.alpha = 0;
                st.beta = 0;
                st.omega = za;  // zero-RHS flag travels in omega (unused by CG)
                st.active = za ? 0 : 1;
                st.converged = za;
                st.finite = 1;
                st.used = -1;
                st.final_res = 0;
                st.best_crit = 0;
                st.last_crit = 0;
                st.best_it = -1;
                st.rising = 0;
                st.save_best = 0;
                st.restore_best = 0;
            }
        });
        zero_all = sst[0].omega != 0.0;
    }
    if (zero_all) {
        for (size_t i = i0; i < (size_t)n; i += stride)
            for (int k = 0; k < NRHS; ++k) a.x[(size_t)k * n + i] = 0;
        if (blockIdx.x == 0 && threadIdx.x == 0)
            for (int k = 0; k < NRHS; ++k) a.st[k] = sst[k];
        return;
    }

    // w = A r (buffer `buf`); gamma = (r,r), delta = (w,r); alpha = gamma/delta, beta = 0
#define PICT_CG1_W_PHASE(buf)                                                                     \
    {                                                                                             \
        bool act[NRHS];                                                                        \
        _Pragma("unroll") for (int k = 0; k < NRHS; ++k) act[k] = k < NRHS && sst[k].active; \
        double acc[NRHS][5];                                                                   \
        zero_acc(acc);                                                                            \
        for (size_t i = i0; i < (size_t)n; i += stride) {                                         \
            const int rs = a.row[i], re = a.row[i + 1];                                           \
            _Pragma("unroll") for (int k = 0; k < NRHS; ++k) {                                 \
                if (!act[k]) continue;                                                            \
                const size_t o = (size_t)k * n;                                                   \
                R w = 0;                                                                          \
                for (int e = rs; e < re; ++e) w += mat_val(a, e) * rb[buf][o + a.col[e]];         \
                wb[buf][o + i] = w;                                                               \
                const R rv = rb[buf][o + i];                                                      \
                acc_res(acc, k, rv);                                                              \
                acc[k][3] += (double)w * (double)rv;                                              \
            }                                                                                     \
        }                                                                                         \
        PICT_FLAT_BARRIER(5, acc, IDS_G, MAX_G, {                                                 \
            for (int k = 0; k < NRHS; ++k) {                                                    \
                RhsState &st = sst[k];                                                            \
                if (!st.active) continue;                                                         \
                st.rho = red[k * kSlots + S_RR];                                                  \
                st.alpha = st.rho / red[k * kSlots + S_D0];                                       \
                st.beta = 0;                                                                      \
            }                                                                                     \
        });                                                                                       \
    }

    PICT_CG1_W_PHASE(0)

    // matrix rows of this thread in registers when every thread owns at most kRowsCached rows
    RowCache<R, kRowsCached> rc;
    // rc.len[q]: -1 no row, -2 row longer than kMaxNz (read from memory), else cached length
    const bool cached = stride * kRowsCached >= (size_t)n;
    if (cached) {
#pragma unroll
        for (int q = 0; q < kRowsCached; ++q) {
            const size_t i = i0 + (size_t)q * stride;
            rc.len[q] = -1;
            if (i < (size_t)n) {
                const int rs = a.row[i], re = a.row[i + 1];
                rc.len[q] = re - rs <= kMaxNz ? re - rs : -2;
#pragma unroll
                for (int e = 0; e < kMaxNz; ++e) {
                    rc.col[q][e] = e < re - rs ? a.col[rs + e] : 0;
                    rc.val[q][e] = e < re - rs ? mat_val(a, rs + e) : R(0);
                }
            }
        }
    }

    for (int it = 0; it < a.maxit; ++it) {
        if (a.reset_steps > 0 && (it + 1) % a.reset_steps == 0) {
            // residual reset (cg_solver_kernel.cu:266-287): r = b - A x, restart with p = r
            {
                bool act[NRHS];
#pragma unroll
                for (int k = 0; k < NRHS; ++k) act[k] = k < NRHS && sst[k].active;
                double acc[NRHS][4];
                zero_acc(acc);
                for (size_t i = i0; i < (size_t)n; i += stride) {
                    const int rs = a.row[i], re = a.row[i + 1];
#pragma unroll
                    for (int k = 0; k < NRHS; ++k) {
                        if (!act[k]) continue;
                        const size_t o = (size_t)k * n;
                        R s = 0;
                        for (int e = rs; e < re; ++e) s += mat_val(a, e) * a.x[o + a.col[e]];
                        rb[cur][o + i] = a.b[o + i] - s;
                        p[o + i] = 0;
                        sb[cur][o + i] = 0;
                    }
                }
                PICT_FLAT_BARRIER(4, acc, IDS_R, MAX_R, {});
            }
            PICT_CG1_W_PHASE(cur)
        }
        const int nxt = cur ^ 1;
        // ---- the iteration: one pass, one reduce-barrier ----
        {
            bool act[NRHS], save[NRHS];
            R alpha[NRHS], beta[NRHS];
#pragma unroll
            for (int k = 0; k < NRHS; ++k) {
                act[k] = k < NRHS && sst[k].active;
                save[k] = act[k] && sst[k].save_best;
                alpha[k] = act[k] ? (R)sst[k].alpha : R(0);
                beta[k] = act[k] ? (R)sst[k].beta : R(0);
            }
            const R *r = rb[cur], *w = wb[cur], *so = sb[cur];
            R *rn = rb[nxt], *wn = wb[nxt], *sn = sb[nxt];
            double acc[NRHS][5];
            zero_acc(acc);
            // one row: update p, s, x, r and form w' = A r' with r' of the neighbours on the fly;
            // `ent(e, c, v)` yields entry e of the row (column c, value v), nz entries
            auto row_update = [&](size_t i, int nz, auto ent) {
#pragma unroll
                for (int k = 0; k < NRHS; ++k) {
                    if (!act[k]) continue;
                    const size_t o = (size_t)k * n;
                    if (save[k]) a.best[o + i] = a.x[o + i];
                    const R pm = fma(beta[k], p[o + i], r[o + i]);
                    const R sm = fma(beta[k], so[o + i], w[o + i]);
                    p[o + i] = pm;
                    sn[o + i] = sm;
                    a.x[o + i] += alpha[k] * pm;
                    const R rm = fma(-alpha[k], sm, r[o + i]);
                    rn[o + i] = rm;
                    R wv = 0;
                    ent(nz, [&](int cidx, R v) {
                        const size_t c = o + cidx;
                        wv += v * fma(-alpha[k], fma(beta[k], so[c], w[c]), r[c]);
                    });
                    wn[o + i] = wv;
                    acc_res(acc, k, rm);
                    acc[k][3] += (double)wv * (double)rm;
                }
            };
            auto mem_row = [&](size_t i) {
                const int rs = a.row[i], re = a.row[i + 1];
                row_update(i, re - rs, [&](int, auto f) {
                    for (int e = rs; e < re; ++e) f(a.col[e], mat_val(a, e));
                });
            };
            // register budget: the gathered form for up to 8 bytes of right-hand-side values per row
            constexpr bool kGather = NRHS * sizeof(R) <= 8;
            if (cached && !kGather) {
#pragma unroll
                for (int q = 0; q < kRowsCached; ++q) {
                    const size_t i = i0 + (size_t)q * stride;
                    if (rc.len[q] >= 0) {
                        row_update(i, rc.len[q], [&](int nz, auto f) {
#pragma unroll
                            for (int e = 0; e < kMaxNz; ++e)
                                if (e < nz) f(rc.col[q][e], rc.val[q][e]);
                        });
                    } else if (rc.len[q] == -2) {
                        mem_row(i);
                    }
                }
            } else if (cached) {
                // gather every load of this thread's cached rows before any store (one memory round trip),
                // then update and store
                R gr[kRowsCached][NRHS][kMaxNz], gw[kRowsCached][NRHS][kMaxNz], gs[kRowsCached][NRHS][kMaxNz];
                R op[kRowsCached][NRHS], ox[kRowsCached][NRHS], ori[kRowsCached][NRHS], owi[kRowsCached][NRHS],
                    osi[kRowsCached][NRHS];
#pragma unroll
                for (int q = 0; q < kRowsCached; ++q) {
                    const size_t i = i0 + (size_t)q * stride;
                    if (rc.len[q] < 0) continue;
#pragma unroll
                    for (int k = 0; k < NRHS; ++k) {
                        if (!act[k]) continue;
                        const size_t o = (size_t)k * n;
                        op[q][k] = p[o + i];
                        ox[q][k] = a.x[o + i];
                        ori[q][k] = r[o + i];
                        owi[q][k] = w[o + i];
                        osi[q][k] = so[o + i];
#pragma unroll
                        for (int e = 0; e < kMaxNz; ++e) {
                            if (e < rc.len[q]) {
                                const size_t c = o + rc.col[q][e];
                                gr[q][k][e] = r[c];
                                gw[q][k][e] = w[c];
                                gs[q][k][e] = so[c];
                            }
                        }
                    }
                }
#pragma unroll
                for (int q = 0; q < kRowsCached; ++q) {
                    const size_t i = i0 + (size_t)q * stride;
                    if (rc.len[q] == -2) {
                        mem_row(i);
                        continue;
                    }
                    if (rc.len[q] < 0) continue;
#pragma unroll
                    for (int k = 0; k < NRHS; ++k) {
                        if (!act[k]) continue;
                        const size_t o = (size_t)k * n;
                        if (save[k]) a.best[o + i] = ox[q][k];
                        const R pm = fma(beta[k], op[q][k], ori[q][k]);
                        const R sm = fma(beta[k], osi[q][k], owi[q][k]);
                        p[o + i] = pm;
                        sn[o + i] = sm;
                        R xv = ox[q][k];
                        xv += alpha[k] * pm;
                        a.x[o + i] = xv;
                        const R rm = fma(-alpha[k], sm, ori[q][k]);
                        rn[o + i] = rm;
                        R wv = 0;
#pragma unroll
                        for (int e = 0; e < kMaxNz; ++e)
                            if (e < rc.len[q]) wv += rc.val[q][e] * fma(-alpha[k], fma(beta[k], gs[q][k][e], gw[q][k][e]), gr[q][k][e]);
                        wn[o + i] = wv;
                        acc_res(acc, k, rm);
                        acc[k][3] += (double)wv * (double)rm;
                    }
                }
            } else {
                for (size_t i = i0; i < (size_t)n; i += stride) mem_row(i);
            }
            PICT_FLAT_BARRIER(5, acc, IDS_G, MAX_G, {
                for (int k = 0; k < NRHS; ++k) {
                    RhsState &st = sst[k];
                    if (!st.active) continue;
                    st.save_best = 0;
                    const double gam = red[k * kSlots + S_RR];
                    const double crit = criterion(a.conv, gam, red[k * kSlots + S_ASUM], red[k * kSlots + S_AMAX], n);
                    if (!isfinite(crit)) {
                        st.converged = 0; st.used = it; st.final_res = crit; st.finite = 0; st.active = 0;
                        continue;
                    }
                    if (a.return_best) {
                        if (it == 0 || crit < st.best_crit) { st.best_crit = crit; st.best_it = it; st.save_best = 1; }
                        if (it > 0 && crit >= st.last_crit) ++st.rising; else st.rising = 0;
                        st.last_crit = crit;
                    }
                    st.used = it;
                    st.final_res = crit;
                    if (crit < a.tol) { st.converged = 1; st.finite = 1; st.active = 0; continue; }
                    if (a.return_best && (it == a.maxit - 1 || st.rising >= 100)) {
                        st.restore_best = st.best_it != it;
                        st.converged = 0; st.used = st.best_it; st.final_res = st.best_crit; st.finite = 1;
                        st.active = 0;
                        continue;
                    }
                    const double beta_n = gam / st.rho;
                    st.alpha = gam / (red[k * kSlots + S_D0] - beta_n * gam / st.alpha);
                    st.beta = beta_n;
                    st.rho = gam;
                }
            });
        }
        cur = nxt;
        bool any = false;
        for (int k = 0; k < NRHS; ++k) any = any || sst[k].active;
        if (!any) break;
    }
#undef PICT_CG1_W_PHASE
    for (int k = 0; k < NRHS; ++k) {
        if (!sst[k].restore_best) continue;
        const size_t o = (size_t)k * n;
        for (size_t i = i0; i < (size_t)n; i += stride) a.x[o + i] = a.best[o + i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (int k = 0; k < NRHS; ++k) a.st[k] = sst[k];
}

// ==========================================================================
// BiCGStab without preconditioner (bicgstabSolveGPU, bicgstab_solver_kernel.cu:222-365)
// Three reduce-barriers per iteration: (rhat,v) | half-step norm, (t,s), (t,t) |
// norm, (rhat,r).  p = r + beta (p - omega v) and s = r - alpha v are folded
// into the following SpMVs, evaluated on the fly for the neighbours.
// ==========================================================================
template <typename R, int NRHS>
__global__ __launch_bounds__(kSolveThreads) void k_bicgstab(SolveArgs<R> a) {
    __shared__ double red[NRHS * kSlots];
    __shared__ RhsState sst[NRHS];
    __shared__ int s_err;
    const int n = a.n;
    const size_t stride = (size_t)a.nwg * kSolveThreads;
    const size_t i0 = (size_t)blockIdx.x * kSolveThreads + threadIdx.x;
    unsigned gen = 0;
    constexpr int IDS_R[4] = {S_RR, S_ASUM, S_AMAX, S_BMAX};
    constexpr bool MAX_R[4] = {false, false, true, true};
    constexpr int IDS_D[1] = {S_D0};
    constexpr bool MAX_D[1] = {false};
    constexpr int IDS_B[5] = {S_RR, S_ASUM, S_AMAX, S_D0, S_D1};
    constexpr bool MAX_B[5] = {false, false, true, false, false};
    constexpr int IDS_C[4] = {S_RR, S_ASUM, S_AMAX, S_D0};
    constexpr bool MAX_C[4] = {false, false, true, false};
    int zero_all = 0;

    // ---- phase 0: r = b - A x, rhat = r, p = v = 0 ----
    {
        double acc[NRHS][4];
        zero_acc(acc);
        for (size_t i = i0; i < (size_t)n; i += stride) {
            const int rs = a.row[i], re = a.row[i + 1];
#pragma unroll
            for (int k = 0; k < NRHS; ++k) {
                
                const size_t o = (size_t)k * n;
                R s = 0;
                for (int e = rs; e < re; ++e) s += mat_val(a, e) * a.x[o + a.col[e]];
                const R bv = a.b[o + i];
                const R rv = bv - s;
                a.r[o + i] = rv;
                a.rhat[o + i] = rv;
                a.p0[o + i] = 0;
                a.v0[o + i] = 0;
                acc_res(acc, k, rv);
                acc[k][3] = fmax(acc[k][3], fabs((double)bv));
            }
        }
        PICT_FLAT_BARRIER(4, acc, IDS_R, MAX_R, {
            bool all_zero = true;
            for (int k = 0; k < NRHS; ++k) all_zero = all_zero && red[k * kSlots + S_BMAX] == 0.0;
            const int za = (a.zero_shortcut && all_zero) ? 1 : 0;
            for (int k = 0; k < NRHS; ++k) {
                RhsState &st = sst[k];
                const double nr0 = criterion(a.conv, red[k * kSlots + S_RR], red[k * kSlots + S_ASUM],
                                             red[k * kSlots + S_AMAX], n);
                st.finite = 1;
                st.restore_best = 0;
                st.save_best = 0;
                st.rho = red[k * kSlots + S_RR];  // (rhat, r) with rhat = r
                st.beta = 0;
                st.alpha = 0;
                st.omega = 0;
                st.used = -1;
                st.best_crit = za;  // zero-RHS flag travels in best_crit (unused by BiCGStab)
                if (za) {
                    st.active = 0; st.converged = 1; st.final_res = 0;
                } else if (nr0 < a.tol) {  // bicgstab_solver_kernel.cu:259, :351-356
                    st.active = 0; st.converged = 1; st.final_res = nr0;
                } else {
                    st.active = 1; st.converged = 0; st.final_res = 0;
                }
            }
        });
        zero_all = sst[0].best_crit != 0.0;
    }
    if (zero_all) {
        for (size_t i = i0; i < (size_t)n; i += stride)
            for (int k = 0; k < NRHS; ++k) a.x[(size_t)k * n + i] = 0;
    }

    for (int it = 0; it < a.maxit && !zero_all; ++it) {
        {
            bool any = false;
            for (int k = 0; k < NRHS; ++k) any = any || sst[k].active;
            if (!any) break;
        }
        R *pold = (it & 1) ? a.p1 : a.p0;
        R *pnew = (it & 1) ? a.p0 : a.p1;
        R *vold = (it & 1) ? a.v1 : a.v0;
        R *vnew = (it & 1) ? a.v0 : a.v1;
        // ---- phase A: p = r + beta (p - omega v) on the fly; v = A p; (rhat, v) ----
        {
            bool act[NRHS];
            R beta[NRHS], mom[NRHS];
#pragma unroll
            for (int k = 0; k < NRHS; ++k) {
                act[k] = k < NRHS && sst[k].active;
                beta[k] = act[k] ? (R)sst[k].beta : R(0);
                mom[k] = act[k] ? (R)(-sst[k].omega) : R(0);
            }
            double acc[NRHS][1];
            zero_acc(acc);
            for (size_t i = i0; i < (size_t)n; i += stride) {
                const int rs = a.row[i], re = a.row[i + 1];
#pragma unroll
                for (int k = 0; k < NRHS; ++k) {
                    if (!act[k]) continue;
                    const size_t o = (size_t)k * n;
                    const R pn = fma(beta[k], fma(mom[k], vold[o + i], pold[o + i]), a.r[o + i]);
                    pnew[o + i] = pn;
                    R s = 0;
                    for (int e = rs; e < re; ++e) {
                        const size_t c = o + a.col[e];
                        s += mat_val(a, e) * fma(beta[k], fma(mom[k], vold[c], pold[c]), a.r[c]);
                    }
                    vnew[o + i] = s;
                    acc[k][0] += (double)a.rhat[o + i] * (double)s;
                }
            }
            PICT_FLAT_BARRIER(1, acc, IDS_D, MAX_D, {
                for (int k = 0; k < NRHS; ++k)
                    if (sst[k].active) sst[k].alpha = sst[k].rho / red[k * kSlots + S_D0];
            });
        }
        // ---- phase B: s = r - alpha v; t = A s (on the fly); x += alpha p ----
        {
            bool act[NRHS];
            R alpha[NRHS];
#pragma unroll
            for (int k = 0; k < NRHS; ++k) {
                act[k] = k < NRHS && sst[k].active;
                alpha[k] = act[k] ? (R)sst[k].alpha : R(0);
            }
            double acc[NRHS][5];
            zero_acc(acc);
            for (size_t i = i0; i < (size_t)n; i += stride) {
                const int rs = a.row[i], re = a.row[i + 1];
#pragma unroll
                for (int k = 0; k < NRHS; ++k) {
                    if (!act[k]) continue;
                    const size_t o = (size_t)k * n;
                    const R mal = -alpha[k];
                    const R sv = fma(mal, vnew[o + i], a.r[o + i]);
                    a.s[o + i] = sv;
                    R tv = 0;
                    for (int e = rs; e < re; ++e) {
                        const size_t c = o + a.col[e];
                        tv += mat_val(a, e) * fma(mal, vnew[c], a.r[c]);
                    }
                    a.t[o + i] = tv;
                    a.x[o + i] += alpha[k] * pnew[o + i];
                    acc_res(acc, k, sv);
                    acc[k][3] += (double)tv * (double)sv;
                    acc[k][4] += (double)tv * (double)tv;
                }
            }
            PICT_FLAT_BARRIER(5, acc, IDS_B, MAX_B, {
                for (int k = 0; k < NRHS; ++k) {
                    RhsState &st = sst[k];
                    if (!st.active) continue;
                    const double nr = criterion(a.conv, red[k * kSlots + S_RR], red[k * kSlots + S_ASUM],
                                                red[k * kSlots + S_AMAX], n);
                    if (!isfinite(nr)) {
                        st.converged = 0; st.used = it; st.final_res = nr; st.finite = 0; st.active = 0;
                        continue;
                    }
                    st.used = it;
                    st.final_res = nr;
                    if (nr < a.tol) { st.converged = 1; st.active = 0; continue; }
                    st.omega = red[k * kSlots + S_D0] / red[k * kSlots + S_D1];
                }
            });
        }
        // ---- phase C: x += omega s; r = s - omega t; (rhat, r), convergence ----
        {
            bool act[NRHS];
            R omega[NRHS];
#pragma unroll
            for (int k = 0; k < NRHS; ++k) {
                act[k] = k < NRHS && sst[k].active;
                omega[k] = act[k] ? (R)sst[k].omega : R(0);
            }
            double acc[NRHS][4];
            zero_acc(acc);
            for (size_t i = i0; i < (size_t)n; i += stride) {
#pragma unroll
                for (int k = 0; k < NRHS; ++k) {
                    if (!act[k]) continue;
                    const size_t o = (size_t)k * n;
                    const R sv = a.s[o + i];
                    a.x[o + i] += omega[k] * sv;
                    const R rv = sv - omega[k] * a.t[o + i];
                    a.r[o + i] = rv;
                    acc_res(acc, k, rv);
                    acc[k][3] += (double)a.rhat[o + i] * (double)rv;
                }
            }
            PICT_FLAT_BARRIER(4, acc, IDS_C, MAX_C, {
                for (int k = 0; k < NRHS; ++k) {
                    RhsState &st = sst[k];
                    if (!st.active) continue;
                    const double nr = criterion(a.conv, red[k * kSlots + S_RR], red[k * kSlots + S_ASUM],
                                                red[k * kSlots + S_AMAX], n);
                    st.final_res = nr;
                    if (nr < a.tol) { st.converged = 1; st.used = it + 1; st.active = 0; continue; }
                    st.rho_prev = st.rho;
                    st.rho = red[k * kSlots + S_D0];
                    st.beta = (st.rho / st.rho_prev) * (st.alpha / st.omega);
                }
            });
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (int k = 0; k < NRHS; ++k) a.st[k] = sst[k];
}

// A^T permutation for a structurally symmetric pattern: perm[e] = index of (j,i) for e = (i,j).
__global__ void k_transpose_perm(const int32_t *row, const int32_t *col, int n, int32_t *perm, int *missing) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int e = row[i]; e < row[i + 1]; ++e) {
        const int j = col[e];
        int lo = row[j], hi = row[j + 1] - 1, found = -1;
        while (lo <= hi) {  // rows are sorted by column
            const int mid = (lo + hi) >> 1;
            const int c = col[mid];
            if (c == i) { found = mid; break; }
            if (c < i) lo = mid + 1; else hi = mid - 1;
        }
        perm[e] = found < 0 ? e : found;
        if (found < 0) atomicExch(missing, 1);
    }
}

template <typename R>
__global__ void k_outer(const R *a, const R *b, const int32_t *row, const int32_t *col, int n, R *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const R ai = a[i];
    for (int e = row[i]; e < row[i + 1]; ++e) out[e] = ai * b[col[e]];
}

}  // namespace pict

using namespace pict;

namespace {

thread_local bool g_timing = false;
thread_local double g_last_ms = 0.0;
thread_local hipEvent_t g_ev[2] = {nullptr, nullptr};

// Occupancy (resident workgroups per CU) of a solver kernel on the current device, cached.
pict_status occupancy(const void *kernel, int &occ, int &cus) {
    struct Entry { int dev; const void *k; int occ; int cus; };
    static thread_local std::vector<Entry> cache;
    int dev;
    PICT_HIP_TRY(hipGetDevice(&dev));
    for (const Entry &e : cache)
        if (e.dev == dev && e.k == kernel) {
            occ = e.occ;
            cus = e.cus;
            return PICT_OK;
        }
    PICT_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    PICT_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, kSolveThreads, 0));
    cache.push_back({dev, kernel, occ, cus});
    return PICT_OK;
}

enum class Kind { kCG1, kCGTextbook, kBiCG };

template <typename R, int NR>
const void *kernel_for(Kind kind) {
    switch (kind) {
        case Kind::kBiCG: return (const void *)k_bicgstab<R, NR>;
        case Kind::kCGTextbook: return (const void *)k_cg<R, NR>;
        default: return (const void *)k_cg1<R, NR>;
    }
}
template <typename R>
const void *kernel_for_nrhs(Kind kind, int nrhs) {
    switch (nrhs) {
        case 1: return kernel_for<R, 1>(kind);
        case 2: return kernel_for<R, 2>(kind);
        case 3: return kernel_for<R, 3>(kind);
        default: return kernel_for<R, 4>(kind);
    }
}

// Queues one solve on `stream`.  When it has run, the final state of each right-hand side is in
// state_out[0..nrhs) and the barrier error flag in *err_out (device memory); ev0/ev1 (optional) bracket
// the solver kernel.
template <typename R>
pict_status enqueue_typed(const pict_csr_desc *A, const void *rhs, void *x, int nrhs, const pict_solve_opts *o,
                          RhsState *state_out, int *err_out, hipEvent_t ev0, hipEvent_t ev1, hipStream_t stream) {
    const int n = A->n;
    // PICT_CG_TEXTBOOK=1 selects the two-barrier textbook CG kernel (diagnostics / parity checks)
    const char *tb_env = getenv("PICT_CG_TEXTBOOK");
    const bool textbook_cg = tb_env && atoi(tb_env) != 0;
    const Kind kind = o->use_bicg ? Kind::kBiCG : (textbook_cg ? Kind::kCGTextbook : Kind::kCG1);
    const void *kernel = kernel_for_nrhs<R>(kind, nrhs);
    int occ = 0, cus = 0;
    pict_status st = occupancy(kernel, occ, cus);
    if (st) return st;
    if (occ < 1) {
        set_error("solver kernel cannot be resident (occupancy 0)");
        return PICT_ERR_HIP;
    }
    // rows per thread: more rows per workgroup means fewer barrier participants
    static const int rows_per_thread = [] {
        const char *e = getenv("PICT_SOLVER_ROWS_PER_THREAD");
        const int v = e ? atoi(e) : 2;
        return v < 1 ? 1 : v;
    }();
    const long long per_wg = (long long)kSolveThreads * rows_per_thread;
    long long nwg = ((long long)n + per_wg - 1) / per_wg;
    const long long cap = (long long)occ * cus;
    if (nwg > cap) nwg = cap;
    if (nwg > 2048) nwg = 2048;
    if (nwg < 1) nwg = 1;

    const size_t vec = (size_t)n * nrhs;
    const int nvec = kind == Kind::kCGTextbook ? (o->return_best_result ? 5 : 4) : 9;
    const size_t bytes_vec = sizeof(R) * vec * nvec;
    const size_t bytes_part = sizeof(double) * (size_t)2 * nwg * kPartStride;
    const size_t bytes_perm = (o->transpose_a && !A->transpose_perm) ? sizeof(int32_t) * (size_t)A->nnz : 0;
    const size_t off_part = (bytes_vec + 255) & ~size_t(255);
    const size_t off_ctl = off_part + ((bytes_part + 255) & ~size_t(255));
    const size_t off_st = off_ctl + ((sizeof(Ctl) + 255) & ~size_t(255));
    const size_t off_perm = off_st + ((sizeof(RhsState) * nrhs + 255) & ~size_t(255));
    const size_t total = off_perm + bytes_perm + 256;

    char *ws = nullptr;
    PICT_HIP_TRY(hipMallocAsync((void **)&ws, total, stream));
    struct Guard {
        char *p; hipStream_t s;
        ~Guard() { if (p) (void)hipFreeAsync(p, s); }
    } guard{ws, stream};

    SolveArgs<R> a;
    memset(&a, 0, sizeof a);
    a.row = A->row;
    a.col = A->index;
    a.val = (const R *)A->value;
    a.b = (const R *)rhs;
    a.x = (R *)x;
    R *v = (R *)ws;
    a.r = v; v += vec;
    a.p0 = v; v += vec;
    a.p1 = v; v += vec;
    a.v0 = v; v += vec;
    if (kind != Kind::kCGTextbook) {
        a.rhat = v; v += vec;
        a.v1 = v; v += vec;
        a.s = v; v += vec;
        a.t = v; v += vec;
        a.best = v; v += vec;
    } else if (o->return_best_result) {
        a.best = v; v += vec;
    }
    a.partials = (double *)(ws + off_part);
    a.ctl = (Ctl *)(ws + off_ctl);
    a.st = (RhsState *)(ws + off_st);
    a.n = n;
    a.nrhs = nrhs;
    a.maxit = o->max_iterations;
    a.conv = o->convergence_criterion;
    a.reset_steps = o->residual_reset_steps;
    a.return_best = o->return_best_result;
    a.zero_shortcut = o->zero_rhs_shortcut;
    a.tol = o->tolerance;
    a.nwg = (unsigned)nwg;
    PICT_HIP_TRY(hipMemsetAsync(ws + off_ctl, 0, sizeof(Ctl), stream));
    if (o->transpose_a && A->transpose_perm) {
        a.perm = A->transpose_perm;
    } else if (o->transpose_a) {
        int32_t *perm = (int32_t *)(ws + off_perm);
        int *missing = &a.ctl->zero_all;
        hipLaunchKernelGGL(k_transpose_perm, dim3((n + 255) / 256), dim3(256), 0, stream, A->row, A->index, n, perm,
                           missing);
        PICT_HIP_TRY(hipGetLastError());
        int h_missing = 0;
        PICT_HIP_TRY(hipMemcpyAsync(&h_missing, missing, sizeof(int), hipMemcpyDeviceToHost, stream));
        PICT_HIP_TRY(hipStreamSynchronize(stream));
        if (h_missing) {
            set_error("transposed solve requires a structurally symmetric CSR pattern");
            return PICT_ERR_UNSUPPORTED;
        }
        a.perm = perm;
    }
    void *args[] = {&a};
    if (ev0) PICT_HIP_TRY(hipEventRecord(ev0, stream));
    PICT_HIP_TRY(hipLaunchCooperativeKernel(kernel, dim3((unsigned)nwg), dim3(kSolveThreads), args, 0, stream));
    if (ev1) PICT_HIP_TRY(hipEventRecord(ev1, stream));
    PICT_HIP_TRY(hipMemcpyAsync(state_out, a.st, sizeof(RhsState) * nrhs, hipMemcpyDeviceToDevice, stream));
    PICT_HIP_TRY(hipMemcpyAsync(err_out, &a.ctl->err, sizeof(int), hipMemcpyDeviceToDevice, stream));
    return PICT_OK;
}

pict_status validate_solve(const pict_csr_desc *A, int nrhs, const pict_solve_opts *opts) {
    PICT_CHECK_ARG(A && opts, "NULL argument");
    PICT_CHECK_ARG(A->n > 0 && A->row && A->index && A->value, "invalid matrix");
    PICT_CHECK_ARG(nrhs >= 1 && nrhs <= kMaxRhs, "nrhs must be 1..%d (got %d)", kMaxRhs, nrhs);
    PICT_CHECK_ARG(opts->max_iterations >= 0, "max_iterations must be >= 0");
    PICT_CHECK_ARG(opts->convergence_criterion >= 0 && opts->convergence_criterion <= 4, "unknown criterion");
    PICT_CHECK_ARG(A->dtype == PICT_DTYPE_F32 || A->dtype == PICT_DTYPE_F64, "unknown dtype %d", A->dtype);
    if (opts->matrix_rank_deficient) {
        set_error("matrixRankDeficient is not implemented by libpict_hip");
        return PICT_ERR_UNSUPPORTED;
    }
    if (opts->use_bicg && opts->bicg_with_preconditioner) {
        set_error("ILU(0)-preconditioned BiCGStab is not implemented by libpict_hip");
        return PICT_ERR_UNSUPPORTED;
    }
    return PICT_OK;
}

}  // namespace

namespace pict {

size_t solve_state_bytes() { return sizeof(RhsState); }

pict_status solve_enqueue(const pict_csr_desc *A, const void *rhs, void *x, int nrhs, const pict_solve_opts *o,
                          void *state_out, int *err_out, hipEvent_t ev0, hipEvent_t ev1, hipStream_t stream) {
    pict_status st = validate_solve(A, nrhs, o);
    if (st) return st;
    if (A->dtype == PICT_DTYPE_F32)
        return enqueue_typed<float>(A, rhs, x, nrhs, o, (RhsState *)state_out, err_out, ev0, ev1, stream);
    return enqueue_typed<double>(A, rhs, x, nrhs, o, (RhsState *)state_out, err_out, ev0, ev1, stream);
}

void solve_state_to_info(const void *states, int nrhs, pict_solve_info *out) {
    const RhsState *hs = (const RhsState *)states;
    for (int k = 0; k < nrhs; ++k) {
        out[k].final_residual = hs[k].final_res;
        out[k].used_iterations = hs[k].used;
        out[k].converged = hs[k].converged;
        out[k].is_finite_residual = hs[k].finite;
        out[k].reserved = 0;
    }
}

}  // namespace pict

namespace {

// Synchronous solve: queue it, wait, report the states as pict_solve_info.
pict_status solve_sync(const pict_csr_desc *A, const void *rhs, void *x, int nrhs, const pict_solve_opts *o,
                       pict_solve_info *infos, hipStream_t stream) {
    struct Out {
        RhsState st[kMaxRhs];
        int err;
    };
    Out *dev = nullptr;
    PICT_HIP_TRY(hipMallocAsync((void **)&dev, sizeof(Out), stream));
    struct Guard {
        Out *p; hipStream_t s;
        ~Guard() { if (p) (void)hipFreeAsync(p, s); }
    } guard{dev, stream};
    if (g_timing && !g_ev[0]) {
        PICT_HIP_TRY(hipEventCreate(&g_ev[0]));
        PICT_HIP_TRY(hipEventCreate(&g_ev[1]));
    }
    pict_status st = pict::solve_enqueue(A, rhs, x, nrhs, o, dev->st, &dev->err, g_timing ? g_ev[0] : nullptr,
                                         g_timing ? g_ev[1] : nullptr, stream);
    if (st) return st;
    Out host;
    PICT_HIP_TRY(hipMemcpyAsync(&host, dev, sizeof(Out), hipMemcpyDeviceToHost, stream));
    PICT_HIP_TRY(hipStreamSynchronize(stream));
    if (g_timing) {
        float ms = 0.f;
        PICT_HIP_TRY(hipEventElapsedTime(&ms, g_ev[0], g_ev[1]));
        g_last_ms = ms;
    }
    if (host.err) {
        set_error("solver grid barrier exceeded its spin limit");
        return PICT_ERR_DEVICE_SYNC;
    }
    pict::solve_state_to_info(host.st, nrhs, infos);
    return PICT_OK;
}

}  // namespace

extern "C" {

pict_status pict_solve_linear(const pict_csr_desc *A, const void *rhs, void *x, int32_t nrhs,
                              const pict_solve_opts *opts, pict_solve_info *infos_out, void *stream) {
    PICT_CHECK_ARG(A && rhs && x && opts && infos_out, "NULL argument");
    pict_status st = validate_solve(A, nrhs, opts);
    if (st) return st;
    return solve_sync(A, rhs, x, nrhs, opts, infos_out, (hipStream_t)stream);
}

pict_status pict_set_kernel_timing(int32_t enable) {
    g_timing = enable != 0;
    return PICT_OK;
}
double pict_last_solve_kernel_ms(void) { return g_last_ms; }

pict_status pict_csr_transpose_permutation(const pict_csr_desc *P, int32_t *perm_out, void *stream) {
    PICT_CHECK_ARG(P && P->row && P->index && perm_out, "pattern or output missing");
    PICT_CHECK_ARG(P->n >= 0 && P->nnz >= 0, "invalid pattern size");
    if (P->n == 0) return PICT_OK;
    hipStream_t s = (hipStream_t)stream;
    int *missing = nullptr;
    PICT_HIP_TRY(hipMallocAsync((void **)&missing, sizeof(int), s));
    struct Guard {
        int *p; hipStream_t s;
        ~Guard() { if (p) (void)hipFreeAsync(p, s); }
    } guard{missing, s};
    PICT_HIP_TRY(hipMemsetAsync(missing, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_transpose_perm, dim3((P->n + 255) / 256), dim3(256), 0, s, P->row, P->index, P->n, perm_out,
                       missing);
    PICT_HIP_TRY(hipGetLastError());
    int h_missing = 0;
    PICT_HIP_TRY(hipMemcpyAsync(&h_missing, missing, sizeof(int), hipMemcpyDeviceToHost, s));
    PICT_HIP_TRY(hipStreamSynchronize(s));
    if (h_missing) {
        set_error("transposed solve requires a structurally symmetric CSR pattern");
        return PICT_ERR_UNSUPPORTED;
    }
    return PICT_OK;
}

pict_status pict_sparse_outer_product(const void *a, const void *b, const pict_csr_desc *pat, void *out,
                                      void *stream) {
    PICT_CHECK_ARG(a && b && pat && out && pat->row && pat->index, "NULL argument");
    const int n = pat->n;
    if (pat->dtype == PICT_DTYPE_F32)
        hipLaunchKernelGGL(k_outer<float>, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, (const float *)a,
                           (const float *)b, pat->row, pat->index, n, (float *)out);
    else
        hipLaunchKernelGGL(k_outer<double>, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                           (const double *)a, (const double *)b, pat->row, pat->index, n, (double *)out);
    PICT_HIP_TRY(hipGetLastError());
    return PICT_OK;
}

}  // extern "C"
