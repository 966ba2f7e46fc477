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

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
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
    unsigned group_count[32][32];  // per-group arrivals (a line per group)
};
// Resident workgroups (256 threads) per compute unit that a persistent solver grid is sized for.  Grids of seven
// and eight per unit (1792 and 2048 workgroups: the 1.77 M-row channel pressure system on the lean grid-wide CG
// and on the textbook CG) ended at the barrier's spin limit.  They sat exactly at the occupancy query's maximum,
// which SGPRs set for those kernels (102-106 per wave: 7 waves per SIMD), while grids at a VGPR-set maximum run;
// the counters and the spin budget do not depend on the grid size.  Four per unit stays below the SGPR limit of
// every solver kernel; kernels whose VGPRs allow fewer take the query's figure.
constexpr int kMaxResidentWgsPerCu = 4;
constexpr int kGroups = 8;       // arrival tree fan-in (one group per XCD when blocks are dealt round-robin)

// Loads and stores that one workgroup uses to exchange data with another: agent-scope atomic
// operations (sc1 on gfx950), coherent across the XCDs' L2 caches without L2 writeback or
// invalidation.
template <typename T>
__device__ __forceinline__ void st_coh(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T ld_coh(const T *p) {
    return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// wait until every memory operation this thread issued has completed
__device__ __forceinline__ void ops_done() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

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
    double *partials;                              // partial_doubles(nwg) (grid-wide kernels)
    Ctl *ctl;
    RhsState *st;                                  // [nrhs]
    int n, nrhs, maxit, conv, reset_steps, return_best, zero_shortcut;
    double tol;
    unsigned nwg;
    double crit_scale;                             // criterion factor: 1/sqrt(n) (NORM2_NORMALIZED), 1/n (ABS_MEAN), 1
    unsigned long long *tagged;                    // tagged partials [2][nwg][kSlots], zeroed (k_cg_local packs its
                                                   // NS values per workgroup: [2][nwg][NS])
    void *pub;                                     // k_cg_local: published rows r, w, s [3][2][n], zeroed
    int salt;                                      // k_cg_local: stamp of this launch's foreign-read marks
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
// Two-level reduce-barrier of the grid-wide kernels (single-reduction CG and BiCGStab).  Workgroup b
// belongs to group g = b mod G, G = min(nwg, kTreeGroups); with blocks dealt round-robin over the
// eight XCDs the members of a group share one XCD.
//  1. Every workgroup stores its block-reduced FP64 partials (double buffered by generation parity)
//     and, after a release fence (its vector stores and partials published), counts on its group's
//     line.
//  2. The last arrival of a group (acquire) reduces the group's partials with wave 0 (lane l takes
//     member g + l G, in lane order), stores the group's totals and counts on the top line.
//  3. The last group (acquire) reduces the G group totals in group order, stores the final totals
//     and opens the next generation (release).
//  4. The other workgroups wait for the generation, acquire, and read the final totals.
// Every workgroup holds bitwise identical totals (one fixed summation order) and runs the same scalar
// update on its own LDS copy of the solver state.  A counter sees at most max(nwg / G, G) arrivals
// per generation: with one counter for all, its 1024 serialised atomics and every workgroup's read
// of all 1024 partials cost 34 us per barrier on the 1.77 M-row channel system.
// Partials region: [2][nwg] workgroup lines, then kTreeGroups group lines and one line of totals.
// --------------------------------------------------------------------------
constexpr int kPartStride = kMaxRhs * kSlots;
constexpr int kTreeGroups = 32;
constexpr size_t partial_doubles(long long nwg) { return ((size_t)2 * (size_t)nwg + kTreeGroups + 1) * kPartStride; }

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

// FENCE = false: the caller exchanges its vectors with agent-scope atomics (st_coh / ld_coh); every wave
// waits for its stores before the workgroup arrives, the barrier's own data go the same way, and no L2
// writeback or invalidation is needed.
template <int NR, int NS, bool FENCE = true>
__device__ __forceinline__ void flat_reduce(double *part, Ctl *ctl, unsigned nwg, unsigned bid, unsigned gen,
                                            double (&acc)[NR][NS], const int (&ids)[NS],
                                            const bool (&ismax)[NS], double *red, int &s_err) {
    __shared__ double wsum[kSolveThreads / 64][NR * NS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    auto release = [] {
        if constexpr (FENCE) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        else ops_done();
    };
    auto acquire = [] {
        if constexpr (FENCE) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    };
    auto put = [](double *p, double v) {
        if constexpr (FENCE) *p = v;
        else st_coh(p, v);
    };
    auto get = [](const double *p) {
        if constexpr (FENCE) return *p;
        else return ld_coh(p);
    };
    if constexpr (!FENCE) ops_done();
    // wave totals, then the workgroup total (fixed order)
#pragma unroll
    for (int k = 0; k < NR; ++k)
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const double w = ismax[s] ? wave_reduce_dpp<true>(acc[k][s]) : wave_reduce_dpp<false>(acc[k][s]);
            if (lane == 0) wsum[wave][k * NS + s] = w;
        }
    __syncthreads();
    if (wave != 0) return;
    const unsigned G = nwg < (unsigned)kTreeGroups ? nwg : (unsigned)kTreeGroups;
    const unsigned g = bid % G;
    const unsigned gsize = (nwg - g + G - 1) / G;
    const double *base = part + (size_t)(gen & 1) * nwg * kPartStride;
    double *gpart = part + (size_t)2 * nwg * kPartStride;
    double *fin = gpart + (size_t)kTreeGroups * kPartStride;
    auto combine = [&](int s, double x, double y) { return ismax[s] ? fmax(x, y) : x + y; };
    // 1. this workgroup's partials; arrive at the group's line
    int role = 0;
    if (lane == 0) {
        double *mine = part + ((size_t)(gen & 1) * nwg + bid) * kPartStride;
#pragma unroll
        for (int k = 0; k < NR; ++k)
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                double t = wsum[0][k * NS + s];
#pragma unroll
                for (int w = 1; w < kSolveThreads / 64; ++w) t = combine(s, t, wsum[w][k * NS + s]);
                put(mine + k * kSlots + ids[s], t);
            }
        release();
        const unsigned prev = __hip_atomic_fetch_add(&ctl->group_count[g][0], 1u, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
        if (prev == gsize - 1) {
            __hip_atomic_store(&ctl->group_count[g][0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            role = 1;
        }
    }
    role = __shfl(role, 0, 64);
    if (role) {
        // 2. the group's totals (members in lane order), then the top line
        acquire();
        double v[NR][NS];
#pragma unroll
        for (int k = 0; k < NR; ++k)
#pragma unroll
            for (int s = 0; s < NS; ++s) v[k][s] = 0.0;
        for (unsigned j = lane; j < gsize; j += 64) {
            const double *pw = base + (size_t)(g + j * G) * kPartStride;
#pragma unroll
            for (int k = 0; k < NR; ++k)
#pragma unroll
                for (int s = 0; s < NS; ++s) v[k][s] = combine(s, v[k][s], get(pw + k * kSlots + ids[s]));
        }
        int top_last = 0;
#pragma unroll
        for (int k = 0; k < NR; ++k)
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const double t = ismax[s] ? wave_reduce_dpp<true>(v[k][s]) : wave_reduce_dpp<false>(v[k][s]);
                if (lane == 0) put(gpart + (size_t)g * kPartStride + k * kSlots + ids[s], t);
            }
        if (lane == 0) {
            release();
            const unsigned top = __hip_atomic_fetch_add(&ctl->count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (top == G - 1) {
                __hip_atomic_store(&ctl->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                top_last = 1;
            }
        }
        top_last = __shfl(top_last, 0, 64);
        if (top_last) {
            // 3. the final totals (groups in lane order); open the next generation
            acquire();
#pragma unroll
            for (int k = 0; k < NR; ++k)
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    v[k][s] = (unsigned)lane < G ? get(gpart + (size_t)lane * kPartStride + k * kSlots + ids[s]) : 0.0;
#pragma unroll
            for (int k = 0; k < NR; ++k)
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    const double t = ismax[s] ? wave_reduce_dpp<true>(v[k][s]) : wave_reduce_dpp<false>(v[k][s]);
                    if (lane == 0) put(fin + k * kSlots + ids[s], t);
                }
            if (lane == 0) {
                release();
                __hip_atomic_fetch_add(&ctl->gen, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
    // 4. wait for the generation to move (relaxed loads; one acquire after), then the totals
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
    acquire();
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NR; ++k)
#pragma unroll
            for (int s = 0; s < NS; ++s) red[k * kSlots + ids[s]] = get(fin + k * kSlots + ids[s]);
        s_err = err;
    }
}

// Reduce-barrier: flat_reduce leaves the totals and the error flag in LDS (written by
// thread 0, which then runs the scalar update); one workgroup barrier publishes both.
#define PICT_FLAT_BARRIER(NS, ACC, IDS, ISMAX, UPDATE) PICT_FLAT_BARRIER_F(NS, ACC, IDS, ISMAX, true, UPDATE)
#define PICT_FLAT_BARRIER_F(NS, ACC, IDS, ISMAX, FENCE, UPDATE)                                   \
    do {                                                                                          \
        flat_reduce<NRHS, NS, FENCE>(a.partials, a.ctl, a.nwg, bid, gen, ACC, IDS, ISMAX, red, s_err); \
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

// One row's entries in registers, loaded back to back (one memory round trip instead of one per
// entry).  nz < 0: the row is empty or longer than kMaxNz (the caller reads it from memory).  Entries
// past nz repeat the last one's column (a valid address to load from) and are never used.
template <typename R>
struct RowRegs {
    int nz;
    int col[kMaxNz];
    R val[kMaxNz];
};
template <typename R>
__device__ __forceinline__ void load_row(const SolveArgs<R> &a, size_t i, RowRegs<R> &rr) {
    const int rs = a.row[i], re = a.row[i + 1];
    const int nz = re - rs;
    rr.nz = nz > 0 && nz <= kMaxNz ? nz : -1;
    if (rr.nz < 0) return;
#pragma unroll
    for (int e = 0; e < kMaxNz; ++e) {
        const int idx = rs + (e < nz ? e : nz - 1);
        rr.col[e] = a.col[idx];
        rr.val[e] = mat_val(a, idx);
    }
}

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

// Scalar step of the single-reduction CG after the reduction of iteration `it`
// (red = this right-hand side's totals): convergence test (cg_solver_kernel.cu:306-313),
// best-result bookkeeping (:330-346, :399-412), then the next alpha and beta.
template <typename R>
__device__ __forceinline__ void cg1_update(const SolveArgs<R> &a, RhsState &st, const double *red, int it) {
    if (!st.active) return;
    st.save_best = 0;
    const double gam = red[S_RR];
    // criterion (ComputeConvergenceCriterion) with the size factor precomputed on the host
    const double crit = (a.conv == PICT_CONV_NORM2 || a.conv == PICT_CONV_NORM2_NORMALIZED) ? sqrt(gam) * a.crit_scale
                        : a.conv == PICT_CONV_ABS_MAX                                      ? red[S_AMAX]
                                                                                           : red[S_ASUM] * a.crit_scale;
    if (!isfinite(crit)) {
        st.converged = 0; st.used = it; st.final_res = crit; st.finite = 0; st.active = 0;
        return;
    }
    if (a.return_best) {
        if (it == 0 || crit < st.best_crit) { st.best_crit = crit; st.best_it = it; st.save_best = 1; }
        if (it > 0 && crit >= st.last_crit) ++st.rising; else st.rising = 0;
        st.last_crit = crit;
    }
    st.used = it;
    st.final_res = crit;
    if (crit < a.tol) { st.converged = 1; st.finite = 1; st.active = 0; return; }
    if (a.return_best && (it == a.maxit - 1 || st.rising >= 100)) {
        st.restore_best = st.best_it != it;
        st.converged = 0; st.used = st.best_it; st.final_res = st.best_crit; st.finite = 1;
        st.active = 0;
        return;
    }
    const double beta_n = gam / st.rho;
    st.alpha = gam / (red[S_D0] - beta_n * gam / st.alpha);
    st.beta = beta_n;
    st.rho = gam;
}

// first reduction of a solve: rho = (r0, r0); zero-RHS shortcut flag in omega
__device__ __forceinline__ void cg1_init(RhsState &st, double rr, int zero_all) {
    st.rho = rr;
    st.alpha = 0;
    st.beta = 0;
    st.omega = zero_all;
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
// LEAN: without the register row cache and its gathered form, for systems whose
// rows outnumber what the cache holds (every thread then streams its rows from
// memory); the cache's registers would only cut the resident waves.
// ==========================================================================
template <typename R, int NRHS, bool LEAN>
__global__ __launch_bounds__(kSolveThreads) void k_cg1(SolveArgs<R> a) {
    __shared__ double red[NRHS * kSlots];
    __shared__ RhsState sst[NRHS];
    __shared__ int s_err;
    const int n = a.n;
    const unsigned bid = blockIdx.x;
    const size_t stride = (size_t)a.nwg * kSolveThreads;
    const size_t i0 = (size_t)bid * kSolveThreads + threadIdx.x;
    unsigned gen = 0;
    R *rb[2] = {a.r, a.rhat}, *wb[2] = {a.v0, a.v1}, *sb[2] = {a.s, a.t};
    R *p = a.p0;
    constexpr int IDS_R[4] = {S_RR, S_ASUM, S_AMAX, S_BMAX};
    constexpr bool MAX_R[4] = {false, false, true, true};
    constexpr int IDS_G[5] = {S_RR, S_ASUM, S_AMAX, S_D0, S_D1};
    constexpr bool MAX_G[5] = {false, false, true, false, false};
    int cur = 0;  // buffer holding r_i, w_i, s_{i-1}
    int zero_all = 0;
    // rows of this thread: grid-stride (i0, i0 + stride, ...) for the kernels with the row cache; the lean kernel
    // walks a contiguous chunk per workgroup, and the chunks of one XCD lie next to each other (workgroup b runs on
    // XCD b mod 8 under round-robin dispatch), so the neighbours a grid line or plane away are mostly rows of the
    // same workgroup or of workgroups that share its L2 (the placement only affects speed)
    size_t r_lo = i0, r_hi = (size_t)n, r_step = stride;
    if constexpr (LEAN) {
        const size_t nw = a.nwg;
        const size_t chunk = ((size_t)n + nw - 1) / nw;
        const size_t c = nw % 8 == 0 ? (size_t)(bid % 8) * (nw / 8) + bid / 8 : (size_t)bid;
        r_lo = c * chunk + threadIdx.x;
        r_hi = min((size_t)n, (c + 1) * chunk);
        r_step = kSolveThreads;
    }

    // ---- phase I0: r0 = b - A x0; p = s = 0 ----
    {
        double acc[NRHS][4];
        zero_acc(acc);
        for (size_t i = r_lo; i < r_hi; i += r_step) {
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
            for (int k = 0; k < NRHS; ++k) cg1_init(sst[k], red[k * kSlots + S_RR], za);
        });
        zero_all = sst[0].omega != 0.0;
    }
    if (zero_all) {
        for (size_t i = r_lo; i < r_hi; i += r_step)
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
        for (size_t i = r_lo; i < r_hi; i += r_step) {                                         \
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
    RowCache<R, LEAN ? 1 : kRowsCached> rc;
    // rc.len[q]: -1 no row, -2 row longer than kMaxNz (read from memory), else cached length
    const bool cached = !LEAN && stride * kRowsCached >= (size_t)n;
    if constexpr (!LEAN) if (cached) {
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
                for (size_t i = r_lo; i < r_hi; i += r_step) {
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
                    const R ri = ld_coh(r + o + i);
                    const R pm = fma(beta[k], p[o + i], ri);
                    const R sm = fma(beta[k], ld_coh(so + o + i), ld_coh(w + o + i));
                    p[o + i] = pm;
                    st_coh(sn + o + i, sm);
                    a.x[o + i] += alpha[k] * pm;
                    const R rm = fma(-alpha[k], sm, ri);
                    st_coh(rn + o + i, rm);
                    R wv = 0;
                    ent(nz, [&](int cidx, R v) {
                        const size_t c = o + cidx;
                        wv += v * fma(-alpha[k], fma(beta[k], ld_coh(so + c), ld_coh(w + c)), ld_coh(r + c));
                    });
                    st_coh(wn + o + i, wv);
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
            if constexpr (LEAN) {
                // per row: its entries (one round trip), then every gather (one more), then the update
                for (size_t i = r_lo; i < r_hi; i += r_step) {
                    RowRegs<R> rr;
                    load_row(a, i, rr);
                    if (rr.nz < 0) {
                        mem_row(i);
                        continue;
                    }
#pragma unroll
                    for (int k = 0; k < NRHS; ++k) {
                        if (!act[k]) continue;
                        const size_t o = (size_t)k * n;
                        R gr[kMaxNz], gw[kMaxNz], gs[kMaxNz];
#pragma unroll
                        for (int e = 0; e < kMaxNz; ++e) {
                            const size_t c = o + rr.col[e];
                            gr[e] = ld_coh(r + c);
                            gw[e] = ld_coh(w + c);
                            gs[e] = ld_coh(so + c);
                        }
                        const R ri = ld_coh(r + o + i), wi = ld_coh(w + o + i), si = ld_coh(so + o + i);
                        if (save[k]) a.best[o + i] = a.x[o + i];
                        const R pm = fma(beta[k], p[o + i], ri);
                        const R sm = fma(beta[k], si, wi);
                        p[o + i] = pm;
                        st_coh(sn + o + i, sm);
                        a.x[o + i] += alpha[k] * pm;
                        const R rm = fma(-alpha[k], sm, ri);
                        st_coh(rn + o + i, rm);
                        R wv = 0;
#pragma unroll
                        for (int e = 0; e < kMaxNz; ++e)
                            if (e < rr.nz) wv += rr.val[e] * fma(-alpha[k], fma(beta[k], gs[e], gw[e]), gr[e]);
                        st_coh(wn + o + i, wv);
                        acc_res(acc, k, rm);
                        acc[k][3] += (double)wv * (double)rm;
                    }
                }
            } else if (cached && !kGather) {
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
                        ori[q][k] = ld_coh(r + o + i);
                        owi[q][k] = ld_coh(w + o + i);
                        osi[q][k] = ld_coh(so + o + i);
#pragma unroll
                        for (int e = 0; e < kMaxNz; ++e) {
                            if (e < rc.len[q]) {
                                const size_t c = o + rc.col[q][e];
                                gr[q][k][e] = ld_coh(r + c);
                                gw[q][k][e] = ld_coh(w + c);
                                gs[q][k][e] = ld_coh(so + c);
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
                        st_coh(sn + o + i, sm);
                        R xv = ox[q][k];
                        xv += alpha[k] * pm;
                        a.x[o + i] = xv;
                        const R rm = fma(-alpha[k], sm, ori[q][k]);
                        st_coh(rn + o + i, rm);
                        R wv = 0;
#pragma unroll
                        for (int e = 0; e < kMaxNz; ++e)
                            if (e < rc.len[q]) wv += rc.val[q][e] * fma(-alpha[k], fma(beta[k], gs[q][k][e], gw[q][k][e]), gr[q][k][e]);
                        st_coh(wn + o + i, wv);
                        acc_res(acc, k, rm);
                        acc[k][3] += (double)wv * (double)rm;
                    }
                }
            } else {
                for (size_t i = r_lo; i < r_hi; i += r_step) mem_row(i);
            }
            // r, w and s travel as agent-scope atomics: the barrier needs no L2 writeback or invalidation,
            // except before a residual reset, which reads the neighbours' x (plain stores)
            if (a.reset_steps > 0) {
                PICT_FLAT_BARRIER(5, acc, IDS_G, MAX_G, {
                    for (int k = 0; k < NRHS; ++k) cg1_update(a, sst[k], red + k * kSlots, it);
                });
            } else {
                PICT_FLAT_BARRIER_F(5, acc, IDS_G, MAX_G, false, {
                    for (int k = 0; k < NRHS; ++k) cg1_update(a, sst[k], red + k * kSlots, it);
                });
            }
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
        for (size_t i = r_lo; i < r_hi; i += r_step) a.x[o + i] = a.best[o + i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (int k = 0; k < NRHS; ++k) a.st[k] = sst[k];
}

// ==========================================================================
// Register-resident single-reduction CG for one right-hand side.
//
// The recurrences of k_cg1 with another placement of the data: workgroup b owns
// the contiguous rows [b*C, min(n, (b+1)*C)), C = 256*RPT, and thread t the rows
// lo + t + q*256 (q < RPT).  Those rows' x, p, r, w, s and matrix entries (up to
// NZ per row; longer rows are read from memory each time) stay in registers for
// the whole solve.  Each iteration r' of the workgroup's own rows goes through
// LDS; for a column owned by another workgroup, r' is formed from the owner's
// published r, w, s of the previous iteration with the owner's fma sequence
// (bitwise equal).  Only rows that another workgroup reads are published (found
// in the first pass: every workgroup stamps the foreign columns it holds), double
// buffered by iteration parity.
//
// Whatever one workgroup writes for another is stored and loaded with
// agent-scope atomic operations, coherent across the XCDs' L2 caches by
// themselves, so no L2 writeback or invalidate is needed; and every such word
// carries the generation it belongs to, so a reader waits on the data itself
// instead of on a separate flag that would have to be ordered after it:
//  - FP64 partial sums carry (generation + 1) mod 4 in the low 2 bits of the
//    mantissa, which readers clear (every workgroup reduces the same truncated
//    values in the same order, so all get bitwise identical totals; truncation
//    changes a partial by < 2^-50 relative, and an exact zero stays zero);
//  - FP32 published rows are 8-byte words (value, iteration + 1).  FP64 rows
//    cannot be packed with a tag in one atomic word: they are plain values and
//    a workgroup waits for its row stores to complete before it stores its
//    partials.
// Both regions are zeroed before the launch (tag 0 is never awaited).
// ==========================================================================

typedef unsigned long long u64;
// A slot is rewritten every second generation, so a stale slot is exactly two generations
// old: (generation + 1) mod 4 tells them apart, and the zeroed start (tag 0) is not awaited
// before every slot has been written once.
__device__ __forceinline__ u64 tag_partial(double v, unsigned gen) {
    return (__double_as_longlong(v) & ~3ull) | (u64)((gen + 1u) & 3u);
}
__device__ __forceinline__ bool partial_ready(u64 u, unsigned gen) { return (unsigned)(u & 3u) == ((gen + 1u) & 3u); }
__device__ __forceinline__ double partial_value(u64 u) { return __longlong_as_double((long long)(u & ~3ull)); }

// Published row values of one vector: [2][n] (parity), FP32 as (value, tag) words.
template <typename R>
struct Pub;
template <>
struct Pub<float> {
    u64 *base;
    int n;
    __device__ __forceinline__ void put(int par, int i, float v, unsigned tag) const {
        st_coh(base + (size_t)par * n + i, ((u64)tag << 32) | (u64)__float_as_uint(v));
    }
    // value of row c (parity par) once it carries `tag`; NaN after a bounded wait
    __device__ __forceinline__ float get(int par, int c, unsigned tag, int &err) const {
        if (err) return __int_as_float(0x7fc00000);
        const u64 *p = base + (size_t)par * n + c;
        u64 u = ld_coh(p);
        unsigned spins = 0;
        while ((unsigned)(u >> 32) != tag) {
            if (++spins > kSpinLimit) {
                err = 1;
                return __int_as_float(0x7fc00000);
            }
            u = ld_coh(p);
        }
        return __uint_as_float((unsigned)u);
    }
    // raw word of row c (one load, no wait), then its value and whether it carries `tag`
    using Word = u64;
    __device__ __forceinline__ u64 word(int par, int c) const { return ld_coh(base + (size_t)par * n + c); }
    static __device__ __forceinline__ float open(u64 u, unsigned tag, bool &ok) {
        ok = ok & ((unsigned)(u >> 32) == tag);
        return __uint_as_float((unsigned)u);
    }
    static constexpr bool kTagged = true;
};
template <>
struct Pub<double> {
    double *base;
    int n;
    __device__ __forceinline__ void put(int par, int i, double v, unsigned) const {
        st_coh(base + (size_t)par * n + i, v);
    }
    __device__ __forceinline__ double get(int par, int c, unsigned, int &) const {
        return ld_coh(base + (size_t)par * n + c);
    }
    using Word = double;
    __device__ __forceinline__ double word(int par, int c) const { return ld_coh(base + (size_t)par * n + c); }
    static __device__ __forceinline__ double open(double u, unsigned, bool &) { return u; }
    static constexpr bool kTagged = false;
};

// Reduce-barrier on tagged partials: the workgroup's totals of acc go to its line of
// the generation's set (tagged); wave 0 then reads all nwg lines until every slot
// carries the generation's tag and reduces them in a fixed order (red[ids[s]], thread
// 0's view).  The two sets alternate by generation parity: a set is written again
// only after every workgroup has passed the next generation.  `wait_rows`: the
// workgroup's row stores must complete before its partials go out (untagged rows).
// Reduction layouts of k_cg_local (slot ids and whether a slot is a maximum), as types so
// that every use folds to constants.  With a norm-2 criterion only (r,r) and (w,r) (and
// max|b| in the first pass) are needed; the sum and the maximum of |r| only for the
// absolute criteria.
struct CgFirstFull {
    static constexpr int NS = 4;
    static constexpr int ID[4] = {S_RR, S_ASUM, S_AMAX, S_BMAX};
    static constexpr bool MAX[4] = {false, false, true, true};
};
struct CgIterFull {
    static constexpr int NS = 4;
    static constexpr int ID[4] = {S_RR, S_ASUM, S_AMAX, S_D0};
    static constexpr bool MAX[4] = {false, false, true, false};
};
struct CgFirst {
    static constexpr int NS = 2;
    static constexpr int ID[2] = {S_RR, S_BMAX};
    static constexpr bool MAX[2] = {false, true};
};
struct CgIter {
    static constexpr int NS = 2;
    static constexpr int ID[2] = {S_RR, S_D0};
    static constexpr bool MAX[2] = {false, false};
};
template <bool FULL>
struct CgSlots {
    using First = CgFirstFull;
    using Iter = CgIterFull;
};
template <>
struct CgSlots<false> {
    using First = CgFirst;
    using Iter = CgIter;
};

// FILL: the workgroup writes every slot of its line (unused ones as 0), for callers whose layouts differ from
// one generation to the next: each slot is then rewritten every second generation, which the 2-bit tag needs
// to tell a stale slot from a current one.
template <class L, bool FILL = false>
__device__ __forceinline__ void tagged_poll(u64 *part, int *err_flag, unsigned nwg, unsigned bid, unsigned gen,
                                            const double (&acc)[L::NS], bool wait_rows, double (&v)[L::NS],
                                            int &s_err) {
    constexpr int NS = L::NS;
    __shared__ double wsum[kSolveThreads / 64][NS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (wait_rows) ops_done();
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const double w = L::MAX[s] ? wave_reduce_dpp<true>(acc[s]) : wave_reduce_dpp<false>(acc[s]);
        if (lane == 0) wsum[wave][s] = w;
    }
    __syncthreads();
    if (wave != 0) return;
    u64 *base = part + (size_t)(gen & 1u) * nwg * kSlots;
    if (lane == 0) {
        u64 *mine = base + (size_t)bid * kSlots;
        double line[kSlots];
#pragma unroll
        for (int j = 0; j < kSlots; ++j) line[j] = 0.0;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            double t = wsum[0][s];
#pragma unroll
            for (int w = 1; w < kSolveThreads / 64; ++w) t = L::MAX[s] ? fmax(t, wsum[w][s]) : t + wsum[w][s];
            if constexpr (FILL) line[L::ID[s]] = t;
            else st_coh(mine + L::ID[s], tag_partial(t, gen));
        }
        if constexpr (FILL) {
#pragma unroll
            for (int j = 0; j < kSlots; ++j) st_coh(mine + j, tag_partial(line[j], gen));
        }
    }
    unsigned spins = 0;
    int err = 0;
    for (;;) {
        bool ready = true;
#pragma unroll
        for (int s = 0; s < NS; ++s) v[s] = 0.0;
        for (unsigned w = lane; w < nwg; w += 64) {
            const u64 *pw = base + (size_t)w * kSlots;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const u64 u = ld_coh(pw + L::ID[s]);
                ready = ready && partial_ready(u, gen);
                const double pv = partial_value(u);
                v[s] = L::MAX[s] ? fmax(v[s], pv) : v[s] + pv;
            }
        }
        if (__all(ready)) break;
        if (++spins > kSpinLimit) {
            err = 1;
            break;
        }
    }
    if (err && lane == 0) st_coh(err_flag, 1);
    if (lane == 0) s_err = err;
}

// the totals of wave 0's polled lines (fixed order) into the slots of red (every lane's own array)
template <class L>
__device__ __forceinline__ void poll_total(const double (&v)[L::NS], double *red) {
#pragma unroll
    for (int s = 0; s < L::NS; ++s)
        red[L::ID[s]] = L::MAX[s] ? wave_reduce_dpp<true>(v[s]) : wave_reduce_dpp<false>(v[s]);
}

template <class L, bool FILL = false>
__device__ __forceinline__ void tagged_reduce(u64 *part, int *err_flag, unsigned nwg, unsigned bid, unsigned gen,
                                              const double (&acc)[L::NS], bool wait_rows, double *red, int &s_err) {
    double v[L::NS];
    tagged_poll<L, FILL>(part, err_flag, nwg, bid, gen, acc, wait_rows, v, s_err);
    if (threadIdx.x < 64) {
        double t[kSlots];
        poll_total<L>(v, t);
        if (threadIdx.x == 0)
#pragma unroll
            for (int s = 0; s < L::NS; ++s) red[L::ID[s]] = t[L::ID[s]];
    }
}

// accumulate the residual terms of one row (slot order of CgSlots)
template <bool FULL, int NS>
__device__ __forceinline__ void cg_acc(double (&acc)[NS], double rv, double last, bool last_is_max) {
    acc[0] += rv * rv;
    if constexpr (FULL) {
        acc[1] += fabs(rv);
        acc[2] = fmax(acc[2], fabs(rv));
    }
    acc[NS - 1] = last_is_max ? fmax(acc[NS - 1], last) : acc[NS - 1] + last;
}

// Reduce-barrier of k_cg_local on tagged partials packed by workgroup: the waves combine their
// totals through LDS, thread 0 stores the workgroup's NS values (tagged) at [set][bid][0..NS), so the
// whole generation's partials lie on a few cache lines (64 workgroups with two values: 1 KB), and
// wave 0 reads them (lane l: workgroups l, l + 64, ..., always in that order) until every value
// carries the generation's tag, so all workgroups form bitwise identical totals.  The two sets
// alternate by generation parity.  `wait_rows`: the rows this workgroup published must have landed
// before its partials go out (untagged FP64 rows).
constexpr int kWaves = kSolveThreads / 64;
// 64-bit words of a tagged partial region (k_cg_local packs NS <= kSlots values per workgroup)
constexpr size_t tagged_words(long long nwg) { return (size_t)2 * (size_t)nwg * kSlots; }

template <class L>
__device__ __forceinline__ void packed_store(u64 *part, unsigned nwg, unsigned bid, unsigned gen, bool wait_rows,
                                             const double (&acc)[L::NS]) {
    constexpr int NS = L::NS;
    __shared__ double wsum[kWaves][NS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (wait_rows) ops_done();
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const double w = L::MAX[s] ? wave_reduce_dpp<true>(acc[s]) : wave_reduce_dpp<false>(acc[s]);
        if (lane == 0) wsum[wave][s] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        u64 *mine = part + ((size_t)(gen & 1u) * nwg + bid) * NS;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            double t = wsum[0][s];
#pragma unroll
            for (int v = 1; v < kWaves; ++v) t = L::MAX[s] ? fmax(t, wsum[v][s]) : t + wsum[v][s];
            st_coh(mine + s, tag_partial(t, gen));
        }
    }
}

// wave 0: the totals of generation gen into t[slot id] (the same value in every lane); false when
// the bounded wait ran out.  Up to 256 workgroups (4 per lane) are read per pass, their loads back
// to back; more in passes of 256.
template <class L>
__device__ __forceinline__ bool packed_poll(const u64 *part, unsigned nwg, unsigned gen, double (&t)[kSlots]) {
    constexpr int NS = L::NS;
    const unsigned lane = threadIdx.x & 63;
    const u64 *set = part + (size_t)(gen & 1u) * nwg * NS;
    double v[NS];
    unsigned spins = 0;
    for (;;) {
        bool ready = true;
#pragma unroll
        for (int s = 0; s < NS; ++s) v[s] = 0.0;
        for (unsigned c0 = 0; c0 < nwg; c0 += 4 * 64) {
            u64 u[4][NS];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const unsigned w = c0 + lane + 64u * j;
                const u64 *pw = set + (size_t)(w < nwg ? w : 0u) * NS;
#pragma unroll
                for (int s = 0; s < NS; ++s) u[j][s] = ld_coh(pw + s);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (c0 + lane + 64u * j >= nwg) continue;
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    ready = ready & partial_ready(u[j][s], gen);
                    const double pv = partial_value(u[j][s]);
                    v[s] = L::MAX[s] ? fmax(v[s], pv) : v[s] + pv;
                }
            }
        }
        if (__all(ready)) break;
        if (++spins > kSpinLimit) return false;
    }
#pragma unroll
    for (int s = 0; s < kSlots; ++s) t[s] = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) t[L::ID[s]] = L::MAX[s] ? wave_reduce_dpp<true>(v[s]) : wave_reduce_dpp<false>(v[s]);
    return true;
}

// 1/d to about one ulp: the hardware estimate and two Newton steps
__device__ __forceinline__ double rcp64(double d) {
    double y = __builtin_amdgcn_rcp(d);
    double e = fma(-d, y, 1.0);
    y = fma(y, e, y);
    e = fma(-d, y, 1.0);
    return fma(y, e, y);
}

// Solver state of k_cg_local's iterations, held by every lane of wave 0 (all lanes compute the
// same values from the same totals).
struct CgRun {
    double rho, inv_rho, den, alpha, beta;  // alpha = rho / den
    double best_crit, last_crit, final_res;
    int active, converged, finite, used, best_it, rising, save_best, restore_best;
};

// cg1_update in the form k_cg_local runs it: the same tests (convergence cg_solver_kernel.cu:306-313,
// best result :330-346), and the next step from
//   beta' = gamma' / rho,  delta' - beta' gamma' / alpha = delta' - beta'^2 den,  alpha' = gamma' / that,
// equal to cg1_update's alpha' and beta' in exact arithmetic.  The candidates are formed before the
// tests, so one reciprocal (not three dependent divisions) lies on the path to the next iteration.
template <bool FULL, typename R>
__device__ __forceinline__ void cg_run_update(const SolveArgs<R> &a, CgRun &st, const double (&t)[kSlots], int it) {
    const double gam = t[S_RR];
    const double beta_n = gam * st.inv_rho;
    const double den_n = fma(-beta_n * beta_n, st.den, t[S_D0]);
    const double alpha_n = gam * rcp64(den_n);
    const double inv_rho_n = rcp64(gam);
    double crit = sqrt(gam) * a.crit_scale;
    if constexpr (FULL)
        crit = (a.conv == PICT_CONV_NORM2 || a.conv == PICT_CONV_NORM2_NORMALIZED) ? crit
               : a.conv == PICT_CONV_ABS_MAX                                      ? t[S_AMAX]
                                                                                  : t[S_ASUM] * a.crit_scale;
    st.save_best = 0;
    if (!isfinite(crit)) {
        st.converged = 0; st.used = it; st.final_res = crit; st.finite = 0; st.active = 0;
        return;
    }
    if (a.return_best) {
        if (it == 0 || crit < st.best_crit) { st.best_crit = crit; st.best_it = it; st.save_best = 1; }
        if (it > 0 && crit >= st.last_crit) ++st.rising; else st.rising = 0;
        st.last_crit = crit;
    }
    st.used = it;
    st.final_res = crit;
    if (crit < a.tol) { st.converged = 1; st.finite = 1; st.active = 0; return; }
    if (a.return_best && (it == a.maxit - 1 || st.rising >= 100)) {
        st.restore_best = st.best_it != it;
        st.converged = 0; st.used = st.best_it; st.final_res = st.best_crit; st.finite = 1;
        st.active = 0;
        return;
    }
    st.alpha = alpha_n;
    st.beta = beta_n;
    st.den = den_n;
    st.rho = gam;
    st.inv_rho = inv_rho_n;
}

// Foreign entries (columns owned by another workgroup) of a thread's rows: at most NH per
// thread, whose owners' words are loaded back to back (unused entries load word 0).  Each
// iteration the thread forms their r' and stores it in its own LDS slots after the workgroup's
// rows (lv[C + t*NH + k]), where the row's register entry points, so every row is one uniform
// gather from LDS.  NH covers the rows of a 2-D grid whose chunk edges cut grid lines (two
// foreign entries per thread) and the z neighbours of a 3-D grid.  Foreign entries beyond a
// thread's NH (a grid whose block connections run along the chunk, such as an O-grid's seam:
// a foreign neighbour in each of a thread's rows) go to the workgroup's shared slots
// lv[C + 256*NH + j], one per thread: thread j loads that column's words and forms its r' with
// the private ones.  A row whose entries fit neither is read from memory like a long row.
template <typename R, int RPT, int NZ, bool FULL>
__device__ __forceinline__ void cg_local(const SolveArgs<R> &a, unsigned bid) {
    constexpr int C = kSolveThreads * RPT;
    constexpr int NH = NZ <= 5 ? 2 : 2 * RPT + 2;
    using LF = typename CgSlots<FULL>::First;
    using LI = typename CgSlots<FULL>::Iter;
    constexpr int NS = LI::NS;
    constexpr bool kTagged = Pub<R>::kTagged;
    __shared__ R lv[C + kSolveThreads * NH + kSolveThreads];  // r (first product) / r' (iterations) of this
                                                             // workgroup's rows, each thread's foreign entries,
                                                             // then the shared foreign slots
    __shared__ int s_ovf_col[kSolveThreads];                 // column of each shared slot (-1: unused)
    __shared__ int s_novf;
    __shared__ double red[kSlots];
    __shared__ RhsState sst;  // solver state through the first two reductions
    __shared__ int s_err;
    // per iteration wave 0 forms the step and publishes what every thread needs
    struct Step {
        double alpha, beta;
        int active, save_best, restore_best;
    };
    __shared__ Step bc;
    const int n = a.n;
    const int lo = (int)bid * C;
    const int cnt = min(C, n - lo);
    unsigned gen = 0;
    using PubT = decltype(Pub<R>::base);
    const Pub<R> pr{reinterpret_cast<PubT>(a.pub), n}, pw{reinterpret_cast<PubT>(a.pub) + 2 * (size_t)n, n},
        ps{reinterpret_cast<PubT>(a.pub) + 4 * (size_t)n, n};
    int *mark = reinterpret_cast<int *>(a.p0);
    const int salt = a.salt;
    int herr = 0;  // a published row did not arrive in time
    auto local = [&](int c) { return (unsigned)(c - lo) < (unsigned)cnt; };
    if (threadIdx.x < kSlots) red[threadIdx.x] = 0.0;
    if (threadIdx.x == 0) {
        s_err = 0;
        s_novf = 0;
    }
    s_ovf_col[threadIdx.x] = -1;
    const int hbase = C + (int)threadIdx.x * NH;  // this thread's LDS slots of foreign entries
    constexpr int obase = C + kSolveThreads * NH;  // the shared slots
    __syncthreads();

    // reduce-barrier of the first two phases: totals into red (thread 0), then UPDATE by thread 0
    auto first_barrier = [&](auto layout, const auto &acc, auto update) {
        using L = decltype(layout);
        packed_store<L>(a.tagged, a.nwg, bid, gen, !kTagged, acc);
        if (threadIdx.x < 64) {
            double t[kSlots];
            const bool ok = packed_poll<L>(a.tagged, a.nwg, gen, t);
            if (threadIdx.x == 0) {
                if (!ok) {
                    st_coh(&a.ctl->err, 1);
                    s_err = 1;
                } else {
#pragma unroll
                    for (int s = 0; s < L::NS; ++s) red[L::ID[s]] = t[L::ID[s]];
                    update();
                }
            }
        }
        ++gen;
        __syncthreads();
    };

    // register rows: matrix entries with their LDS index (own column: c - lo; foreign: the thread's slot;
    // padding: index 0, value 0)
    int col[RPT][NZ];
    R val[RPT][NZ];
    bool row[RPT], mem[RPT], pub[RPT];  // row present / read from memory / read by another workgroup
    R x[RPT], p[RPT], r[RPT], w[RPT], s[RPT], xb[RPT];
    int nh = 0, hc[NH];  // foreign entries: global column

    // ---- rows into registers; stamp the foreign columns; r0 = b - A x0 (published, tag 1) ----
    {
        double acc[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) acc[k] = 0.0;
#pragma unroll
        for (int k = 0; k < NH; ++k) hc[k] = 0;
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int li = (int)threadIdx.x + q * kSolveThreads;
            row[q] = li < cnt;
            mem[q] = false;
            pub[q] = false;
            x[q] = p[q] = r[q] = w[q] = s[q] = xb[q] = R(0);
#pragma unroll
            for (int e = 0; e < NZ; ++e) {
                col[q][e] = 0;
                val[q][e] = R(0);
            }
            if (!row[q]) continue;
            const int i = lo + li;
            const int rs = a.row[i], re = a.row[i + 1];
            int nf = 0;
            for (int e = rs; e < re; ++e) nf += local(a.col[e]) ? 0 : 1;
            // foreign entries: the thread's free private slots first, then shared slots
            const int npriv = min(nf, NH - nh), nshared = nf - npriv;
            int sj = 0;
            if (nshared > 0 && re - rs <= NZ) sj = atomicAdd(&s_novf, nshared);
            mem[q] = re - rs > NZ || sj + nshared > kSolveThreads;
            R sum = 0;
            for (int e = rs; e < re; ++e) {
                const int c = a.col[e];
                const R v = mat_val(a, e);
                if (!mem[q]) {
                    int li_c = c - lo;
                    if (!local(c)) {
                        if (nh < NH) {
                            li_c = hbase + nh;
#pragma unroll
                            for (int k = 0; k < NH; ++k)
                                if (k == nh) hc[k] = c;
                            ++nh;
                        } else {
                            li_c = obase + sj;
                            s_ovf_col[sj] = c;
                            ++sj;
                        }
                    }
#pragma unroll
                    for (int k = 0; k < NZ; ++k)
                        if (k == e - rs) {
                            col[q][k] = li_c;
                            val[q][k] = v;
                        }
                }
                if (!local(c)) st_coh(mark + c, salt);
                sum += v * a.x[c];
            }
            const R bv = a.b[i];
            const R rv = bv - sum;
            x[q] = xb[q] = a.x[i];
            r[q] = rv;
            lv[li] = rv;
            pr.put(0, i, rv, 1u);
            ps.put(0, i, R(0), 1u);
            cg_acc<FULL, NS>(acc, (double)rv, fabs((double)bv), true);
        }
        // the stamps must be visible when the generation completes
        ops_done();
        first_barrier(LF{}, acc, [&] {
            const int za = (a.zero_shortcut && red[S_BMAX] == 0.0) ? 1 : 0;
            cg1_init(sst, red[S_RR], za);
        });
        if (s_err) return;
    }
    if (sst.omega != 0.0) {
#pragma unroll
        for (int q = 0; q < RPT; ++q)
            if (row[q]) a.x[lo + (int)threadIdx.x + q * kSolveThreads] = 0;
        if (bid == 0 && threadIdx.x == 0) a.st[0] = sst;
        return;
    }
#pragma unroll
    for (int q = 0; q < RPT; ++q)
        if (row[q]) pub[q] = ld_coh(mark + lo + (int)threadIdx.x + q * kSolveThreads) == salt;
    const int oc = s_ovf_col[threadIdx.x];  // the column of this thread's shared slot (-1: none)
    const int oslot = obase + (int)threadIdx.x;

    // one row read from memory: sum of A_ij * v(j), v(j) = lv for local columns, else `far(j)`
    auto mem_row = [&](int i, auto far) {
        const int rs = a.row[i], re = a.row[i + 1];
        R acc = 0;
        for (int e = rs; e < re; ++e) {
            const int c = a.col[e];
            acc += mat_val(a, e) * (local(c) ? lv[c - lo] : far(c));
        }
        return acc;
    };
    // one register row: a gather from LDS
    auto product = [&](int q) {
        R wv = 0;
#pragma unroll
        for (int e = 0; e < NZ; ++e) wv += val[q][e] * lv[col[q][e]];
        return wv;
    };

    // ---- w = A r (published, tag 1); gamma = (r,r), delta = (w,r); alpha = gamma/delta ----
    {
        // all loads first (no use in between), then the tag checks; each thread's own LDS slots
        typename Pub<R>::Word ur[NH];
#pragma unroll
        for (int k = 0; k < NH; ++k) ur[k] = pr.word(0, hc[k]);
        R fr[NH];
        bool ok = true;
#pragma unroll
        for (int k = 0; k < NH; ++k) fr[k] = k < nh ? Pub<R>::open(ur[k], 1u, ok) : R(0);
        if (!ok) {
#pragma unroll
            for (int k = 0; k < NH; ++k)
                if (k < nh) fr[k] = pr.get(0, hc[k], 1u, herr);
        }
#pragma unroll
        for (int k = 0; k < NH; ++k) lv[hbase + k] = fr[k];
        if (oc >= 0) lv[oslot] = pr.get(0, oc, 1u, herr);
        __syncthreads();  // the shared slots are read by other threads
        double acc[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) acc[k] = 0.0;
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            if (!row[q]) continue;
            const int i = lo + (int)threadIdx.x + q * kSolveThreads;
            const R wv = !mem[q] ? product(q) : mem_row(i, [&](int c) { return pr.get(0, c, 1u, herr); });
            w[q] = wv;
            if (pub[q]) pw.put(0, i, wv, 1u);
            cg_acc<FULL, NS>(acc, (double)r[q], (double)wv * (double)r[q], false);
        }
        first_barrier(LI{}, acc, [&] {
            if (sst.active) {
                sst.rho = red[S_RR];
                sst.alpha = sst.rho / red[S_D0];
                sst.beta = 0;
            }
        });
        if (s_err) return;
    }

    // the owners' r, w, s of the foreign entries for the next iteration (parity par), all loads back to back
    typename Pub<R>::Word ur[NH], uw[NH], us[NH], uro = {}, uwo = {}, uso = {};
    auto load_foreign = [&](int par) {
#pragma unroll
        for (int k = 0; k < NH; ++k) {
            ur[k] = pr.word(par, hc[k]);
            uw[k] = pw.word(par, hc[k]);
            us[k] = ps.word(par, hc[k]);
        }
        if (oc >= 0) {
            uro = pr.word(par, oc);
            uwo = pw.word(par, oc);
            uso = ps.word(par, oc);
        }
    };
    // FP32 (tagged words): load until every used entry carries `tag` (bounded)
    auto fetch_foreign = [&](int par, unsigned tag) {
        unsigned spins = 0;
        for (;;) {
            load_foreign(par);
            bool ok = true;
#pragma unroll
            for (int k = 0; k < NH; ++k) {
                bool m = true;
                Pub<R>::open(ur[k], tag, m);
                Pub<R>::open(uw[k], tag, m);
                Pub<R>::open(us[k], tag, m);
                ok = ok & (m | (k >= nh));
            }
            {
                bool m = true;
                Pub<R>::open(uro, tag, m);
                Pub<R>::open(uwo, tag, m);
                Pub<R>::open(uso, tag, m);
                ok = ok & (m | (oc < 0));
            }
            if (ok) return;
            if (++spins > kSpinLimit) {
                herr = 1;
                return;
            }
        }
    };
    if constexpr (kTagged) fetch_foreign(0, 1u);
    else load_foreign(0);
    CgRun run{};
    if (threadIdx.x < 64) {
        run.rho = sst.rho;
        run.alpha = sst.alpha;
        run.beta = sst.beta;
        run.den = red[S_D0];
        run.inv_rho = 1.0 / sst.rho;
        run.best_crit = sst.best_crit;
        run.last_crit = sst.last_crit;
        run.final_res = sst.final_res;
        run.active = sst.active;
        run.converged = sst.converged;
        run.finite = sst.finite;
        run.used = sst.used;
        run.best_it = sst.best_it;
        run.rising = sst.rising;
        run.save_best = sst.save_best;
        run.restore_best = sst.restore_best;
    }
    Step cs{sst.alpha, sst.beta, sst.active, sst.save_best, sst.restore_best};
    // a wave whose rows are all present and all in registers runs the iteration without the per-row tests
    bool irregular = false;
#pragma unroll
    for (int q = 0; q < RPT; ++q) irregular = irregular | !row[q] | mem[q];
    const bool simple = !__any(irregular);
    for (int it = 0; it < a.maxit && cs.active; ++it) {
        const int cur = it & 1, nxt = cur ^ 1;
        const unsigned tin = (unsigned)it + 1u, tout = (unsigned)it + 2u;
        const R alpha = (R)cs.alpha, beta = (R)cs.beta;
        const bool save = cs.save_best;
        // own rows: p, s, x, r
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            if (!simple && !row[q]) continue;
            if (save) xb[q] = x[q];
            const R pm = fma(beta, p[q], r[q]);
            const R sm = fma(beta, s[q], w[q]);
            x[q] = fma(alpha, pm, x[q]);
            const R rm = fma(-alpha, sm, r[q]);
            p[q] = pm;
            s[q] = sm;
            r[q] = rm;
            lv[(int)threadIdx.x + q * kSolveThreads] = rm;
        }
        // r' of the foreign entries, the owner's fma sequence, into this thread's LDS slots
        if constexpr (kTagged) {
            bool ok = true;
#pragma unroll
            for (int k = 0; k < NH; ++k) {
                bool m = true;
                Pub<R>::open(ur[k], tin, m);
                Pub<R>::open(uw[k], tin, m);
                Pub<R>::open(us[k], tin, m);
                ok = ok & (m | (k >= nh));
            }
            {
                bool m = true;
                Pub<R>::open(uro, tin, m);
                Pub<R>::open(uwo, tin, m);
                Pub<R>::open(uso, tin, m);
                ok = ok & (m | (oc < 0));
            }
            if (!ok) fetch_foreign(cur, tin);
        }
#pragma unroll
        for (int k = 0; k < NH; ++k) {
            bool m = true;
            const R fr = Pub<R>::open(ur[k], tin, m), fw = Pub<R>::open(uw[k], tin, m),
                    fs = Pub<R>::open(us[k], tin, m);
            lv[hbase + k] = fma(-alpha, fma(beta, fs, fw), fr);
        }
        if (oc >= 0) {
            bool m = true;
            const R fr = Pub<R>::open(uro, tin, m), fw = Pub<R>::open(uwo, tin, m), fs = Pub<R>::open(uso, tin, m);
            lv[oslot] = fma(-alpha, fma(beta, fs, fw), fr);
        }
        __syncthreads();
        // w' = A r'
        double acc[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) acc[k] = 0.0;
        auto finish_row = [&](int q, R wv) {
            const int i = lo + (int)threadIdx.x + q * kSolveThreads;
            w[q] = wv;
            if (pub[q]) {
                pr.put(nxt, i, r[q], tout);
                pw.put(nxt, i, wv, tout);
                ps.put(nxt, i, s[q], tout);
            }
            cg_acc<FULL, NS>(acc, (double)r[q], (double)wv * (double)r[q], false);
        };
        if (simple) {
#pragma unroll
            for (int q = 0; q < RPT; ++q) finish_row(q, product(q));
        } else {
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                if (!row[q]) continue;
                const int i = lo + (int)threadIdx.x + q * kSolveThreads;
                const R wv = !mem[q] ? product(q) : mem_row(i, [&](int c) {
                    return fma(-alpha, fma(beta, ps.get(cur, c, tin, herr), pw.get(cur, c, tin, herr)),
                               pr.get(cur, c, tin, herr));
                });
                finish_row(q, wv);
            }
        }
        // reduce-barrier.  Waves 1.. fetch the neighbours' words as soon as the workgroup's partials are out;
        // wave 0 waits for the totals first (a wave's loads also wait for its earlier stores, and a word
        // fetched before its owner stored it costs another round trip), issues its loads, and forms the step
        // while they travel (checked at the top of the next iteration)
        packed_store<LI>(a.tagged, a.nwg, bid, gen, !kTagged, acc);
        if constexpr (kTagged)
            if (threadIdx.x >= 64) fetch_foreign(nxt, tout);
        if (threadIdx.x < 64) {
            double t[kSlots];
            if (!packed_poll<LI>(a.tagged, a.nwg, gen, t)) {
                if (threadIdx.x == 0) {
                    st_coh(&a.ctl->err, 1);
                    s_err = 1;
                }
            } else {
                if constexpr (kTagged) load_foreign(nxt);
                cg_run_update<FULL>(a, run, t, it);
                if (threadIdx.x == 0) bc = Step{run.alpha, run.beta, run.active, run.save_best, run.restore_best};
            }
        }
        ++gen;
        __syncthreads();
        if (s_err) return;
        // FP64 rows carry no tag: their loads go out once the generation has completed
        if constexpr (!kTagged) load_foreign(nxt);
        cs = bc;
    }
    if (herr) st_coh(&a.ctl->err, 1);
    const bool restore = cs.restore_best;
#pragma unroll
    for (int q = 0; q < RPT; ++q)
        if (row[q]) a.x[lo + (int)threadIdx.x + q * kSolveThreads] = restore ? xb[q] : x[q];
    if (bid == 0 && threadIdx.x == 0) {
        RhsState o = sst;
        o.rho = run.rho;
        o.alpha = run.alpha;
        o.beta = run.beta;
        o.best_crit = run.best_crit;
        o.last_crit = run.last_crit;
        o.final_res = run.final_res;
        o.active = run.active;
        o.converged = run.converged;
        o.finite = run.finite;
        o.used = run.used;
        o.best_it = run.best_it;
        o.rising = run.rising;
        o.save_best = run.save_best;
        o.restore_best = run.restore_best;
        a.st[0] = o;
    }
}

template <typename R, int RPT, int NZ, bool FULL>
__global__ __launch_bounds__(kSolveThreads) void k_cg_local(SolveArgs<R> a) {
    cg_local<R, RPT, NZ, FULL>(a, blockIdx.x);
}

// Batched launch: workgroup w runs jobs[job_off[w] .. job_off[w+1]) in order, job = (system << 16) |
// workgroup index within the system.  Every workgroup's list follows one global order of the systems,
// and the launch is cooperative (all workgroups resident), so the first unfinished system always has
// all its workgroups on it.
template <typename R, int RPT, int NZ, bool FULL>
__global__ __launch_bounds__(kSolveThreads) void k_cg_local_jobs(const SolveArgs<R> *__restrict__ args,
                                                                 const unsigned *__restrict__ job_off,
                                                                 const unsigned *__restrict__ jobs) {
    const unsigned j0 = job_off[blockIdx.x], j1 = job_off[blockIdx.x + 1];
    for (unsigned j = j0; j < j1; ++j) {
        const unsigned job = jobs[j];
        const SolveArgs<R> a = args[job >> 16];
        cg_local<R, RPT, NZ, FULL>(a, job & 0xffffu);
        __syncthreads();
    }
}

// ==========================================================================
// Register-resident BiCGStab (FP32, up to two right-hand sides, norm-2 criteria).
//
// The iteration of k_bicgstab (bicgstabSolveGPU, bicgstab_solver_kernel.cu:222-365) with
// k_cg_local's placement of the data: workgroup b owns rows [b*C, min(n, (b+1)*C)),
// C = 256*RPT; their matrix entries and x, r, rhat, p, v of every right-hand side stay
// in registers for the whole solve.  Each iteration needs two products with foreign
// columns: p (for v = A p) and s (for t = A s).  The owner of a row that another
// workgroup reads publishes p and s as tagged words (value, iteration + 1); a reader
// waits on the word itself.  One buffer per vector suffices: a neighbour reads the
// owner's p (s) before it contributes to the reduce-barrier that follows, and the
// owner writes the next value only after that barrier.  Three reduce-barriers per
// iteration on tagged partials ((rhat, v) | |s|^2, (t, s), (t, t) | |r|^2, (rhat, r)),
// as in k_bicgstab.
// ==========================================================================
template <int NR> struct BicgFirst;
template <int NR> struct BicgA;
template <int NR> struct BicgB;
template <int NR> struct BicgC;
// per right-hand side k: slot 2k = |r|^2, 2k+1 = max|b|
template <> struct BicgFirst<1> { static constexpr int NS = 2; static constexpr int ID[2] = {0, 1}; static constexpr bool MAX[2] = {false, true}; };
template <> struct BicgFirst<2> { static constexpr int NS = 4; static constexpr int ID[4] = {0, 1, 2, 3}; static constexpr bool MAX[4] = {false, true, false, true}; };
// slot k = (rhat, v)
template <> struct BicgA<1> { static constexpr int NS = 1; static constexpr int ID[1] = {0}; static constexpr bool MAX[1] = {false}; };
template <> struct BicgA<2> { static constexpr int NS = 2; static constexpr int ID[2] = {0, 1}; static constexpr bool MAX[2] = {false, false}; };
// slots 3k, 3k+1, 3k+2 = |s|^2, (t, s), (t, t)
template <> struct BicgB<1> { static constexpr int NS = 3; static constexpr int ID[3] = {0, 1, 2}; static constexpr bool MAX[3] = {false, false, false}; };
template <> struct BicgB<2> { static constexpr int NS = 6; static constexpr int ID[6] = {0, 1, 2, 3, 4, 5}; static constexpr bool MAX[6] = {false, false, false, false, false, false}; };
// slots 2k, 2k+1 = |r|^2, (rhat, r)
template <> struct BicgC<1> { static constexpr int NS = 2; static constexpr int ID[2] = {0, 1}; static constexpr bool MAX[2] = {false, false}; };
template <> struct BicgC<2> { static constexpr int NS = 4; static constexpr int ID[4] = {0, 1, 2, 3}; static constexpr bool MAX[4] = {false, false, false, false}; };

template <int RPT, int NZ, int NR>
__device__ __forceinline__ void bicg_local(const SolveArgs<float> &a, unsigned bid) {
    using R = float;
    constexpr int C = kSolveThreads * RPT;
    constexpr int NH = NZ <= 5 ? 2 : 2 * RPT + 2;
    __shared__ R lv[NR][C];  // p, then s, of this workgroup's rows
    __shared__ double red[kSlots];
    __shared__ RhsState sst[NR];
    __shared__ int s_err;
    struct Step {
        R coef[NR];   // beta (phase A), alpha (phase B), omega (phase C)
        R coef2[NR];  // omega of the previous iteration (phase A)
        int active[NR];
    };
    __shared__ Step bc;
    Step cs{};  // this thread's copy of the step coefficients and active flags
    const int n = a.n;
    const int lo = (int)bid * C;
    const int cnt = min(C, n - lo);
    unsigned gen = 0;
    const Pub<R> pp{reinterpret_cast<u64 *>(a.pub), n}, ps{reinterpret_cast<u64 *>(a.pub) + (size_t)NR * n, n};
    int *mark = reinterpret_cast<int *>(a.p0);
    const int salt = a.salt;
    int herr = 0;
    auto local = [&](int c) { return (unsigned)(c - lo) < (unsigned)cnt; };

    int col[RPT][NZ];
    R val[RPT][NZ];
    bool row[RPT], mem[RPT], pub[RPT];
    R x[NR][RPT], r[NR][RPT], rh[NR][RPT], p[NR][RPT], v[NR][RPT];
    int nh = 0, hq[NH], hc[NH];
    R hv[NH];

    // ---- rows into registers; stamp foreign columns; r0 = b - A x0, rhat = r0 ----
    {
        double acc[BicgFirst<NR>::NS];
#pragma unroll
        for (int k = 0; k < BicgFirst<NR>::NS; ++k) acc[k] = 0.0;
#pragma unroll
        for (int k = 0; k < NH; ++k) {
            hq[k] = -1;
            hc[k] = 0;
            hv[k] = R(0);
        }
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int li = (int)threadIdx.x + q * kSolveThreads;
            row[q] = li < cnt;
            mem[q] = false;
            pub[q] = false;
#pragma unroll
            for (int k = 0; k < NR; ++k) x[k][q] = r[k][q] = rh[k][q] = p[k][q] = v[k][q] = R(0);
#pragma unroll
            for (int e = 0; e < NZ; ++e) {
                col[q][e] = lo;
                val[q][e] = R(0);
            }
            if (!row[q]) continue;
            const int i = lo + li;
            const int rs = a.row[i], re = a.row[i + 1];
            int nf = 0;
            for (int e = rs; e < re; ++e) nf += local(a.col[e]) ? 0 : 1;
            mem[q] = re - rs > NZ || nh + nf > NH;
            R sum[NR];
#pragma unroll
            for (int k = 0; k < NR; ++k) sum[k] = 0;
            for (int e = rs; e < re; ++e) {
                const int c = a.col[e];
                const R m = mat_val(a, e);
                if (!mem[q]) {
                    if (local(c)) {
#pragma unroll
                        for (int z = 0; z < NZ; ++z)
                            if (z == e - rs) {
                                col[q][z] = c;
                                val[q][z] = m;
                            }
                    } else {
#pragma unroll
                        for (int z = 0; z < NH; ++z)
                            if (z == nh) {
                                hq[z] = q;
                                hc[z] = c;
                                hv[z] = m;
                            }
                        ++nh;
                    }
                }
                if (!local(c)) st_coh(mark + c, salt);
#pragma unroll
                for (int k = 0; k < NR; ++k) sum[k] += m * a.x[(size_t)k * n + c];
            }
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                const R bv = a.b[(size_t)k * n + i];
                const R rv = bv - sum[k];
                x[k][q] = a.x[(size_t)k * n + i];
                r[k][q] = rh[k][q] = rv;
                acc[2 * k] += (double)rv * (double)rv;
                acc[2 * k + 1] = fmax(acc[2 * k + 1], fabs((double)bv));
            }
        }
        ops_done();  // the stamps before the partials
        if (threadIdx.x < kSlots) red[threadIdx.x] = 0.0;
        tagged_reduce<BicgFirst<NR>, true>(a.tagged, &a.ctl->err, a.nwg, bid, gen, acc, false, red, s_err);
        ++gen;
        if (threadIdx.x == 0 && !s_err) {
            bool all_zero = true;
            for (int k = 0; k < NR; ++k) all_zero = all_zero && red[2 * k + 1] == 0.0;
            const int za = (a.zero_shortcut && all_zero) ? 1 : 0;
            for (int k = 0; k < NR; ++k) {
                RhsState &st = sst[k];
                const double nr0 = criterion(a.conv, red[2 * k], 0.0, 0.0, n);
                st.finite = 1;
                st.restore_best = 0;
                st.save_best = 0;
                st.rho = red[2 * k];
                st.beta = 0;
                st.alpha = 0;
                st.omega = 0;
                st.used = -1;
                st.best_crit = za;
                if (za) {
                    st.active = 0; st.converged = 1; st.final_res = 0;
                } else if (nr0 < a.tol) {
                    st.active = 0; st.converged = 1; st.final_res = nr0;
                } else {
                    st.active = 1; st.converged = 0; st.final_res = 0;
                }
                bc.coef[k] = 0;
                bc.coef2[k] = 0;
                bc.active[k] = st.active;
            }
        }
        __syncthreads();
        if (s_err) return;
    }
    if (sst[0].best_crit != 0.0) {
#pragma unroll
        for (int q = 0; q < RPT; ++q)
            if (row[q])
#pragma unroll
                for (int k = 0; k < NR; ++k) a.x[(size_t)k * n + lo + (int)threadIdx.x + q * kSolveThreads] = 0;
        if (bid == 0 && threadIdx.x < NR) a.st[threadIdx.x] = sst[threadIdx.x];
        return;
    }
#pragma unroll
    for (int q = 0; q < RPT; ++q)
        if (row[q]) pub[q] = ld_coh(mark + lo + (int)threadIdx.x + q * kSolveThreads) == salt;

    // y = A z for the own rows of right-hand side k: z of local columns in LDS, foreign ones from `pv`
    // (published words of the owners, tag `tag`)
    auto product = [&](int k, int q, const Pub<R> &pv, unsigned tag, const R (&fz)[NR][NH]) {
        const int i = lo + (int)threadIdx.x + q * kSolveThreads;
        R y = 0;
        if (!mem[q]) {
#pragma unroll
            for (int e = 0; e < NZ; ++e) y += val[q][e] * lv[k][col[q][e] - lo];
#pragma unroll
            for (int z = 0; z < NH; ++z) y += hq[z] == q ? hv[z] * fz[k][z] : R(0);
        } else {
            const int rs = a.row[i], re = a.row[i + 1];
            for (int e = rs; e < re; ++e) {
                const int c = a.col[e];
                y += mat_val(a, e) * (local(c) ? lv[k][c - lo] : pv.get(0, k * n + c, tag, herr));
            }
        }
        return y;
    };
    // the foreign values of vector `pv` (tag) of the active right-hand sides (only those are published): all
    // loads back to back, then waits for those not yet there
    auto gather = [&](const Pub<R> &pv, unsigned tag, R (&fz)[NR][NH]) {
        u64 w[NR][NH];
#pragma unroll
        for (int k = 0; k < NR; ++k)
#pragma unroll
            for (int z = 0; z < NH; ++z) w[k][z] = pv.word(0, k * n + hc[z]);
        bool ok = true;
#pragma unroll
        for (int k = 0; k < NR; ++k)
#pragma unroll
            for (int z = 0; z < NH; ++z) {
                bool m = true;
                fz[k][z] = Pub<R>::open(w[k][z], tag, m);
                ok = ok & (m | (z >= nh) | !cs.active[k]);
            }
        if (__any(!ok)) {
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                if (!cs.active[k]) continue;
#pragma unroll
                for (int z = 0; z < NH; ++z)
                    if (z < nh) fz[k][z] = pv.get(0, k * n + hc[z], tag, herr);
            }
        }
    };

    cs = bc;
    for (int it = 0; it < a.maxit; ++it) {
        bool any = false;
#pragma unroll
        for (int k = 0; k < NR; ++k) any = any || cs.active[k];
        if (!any) break;
        const unsigned tag = (unsigned)it + 1u;
        // ---- phase A: p = r + beta (p - omega v); v = A p; (rhat, v) ----
        {
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                if (!cs.active[k]) continue;
                const R beta = cs.coef[k], mom = -cs.coef2[k];
#pragma unroll
                for (int q = 0; q < RPT; ++q) {
                    if (!row[q]) continue;
                    const R pn = fma(beta, fma(mom, v[k][q], p[k][q]), r[k][q]);
                    p[k][q] = pn;
                    const int li = (int)threadIdx.x + q * kSolveThreads;
                    lv[k][li] = pn;
                    if (pub[q]) pp.put(0, k * n + lo + li, pn, tag);
                }
            }
            __syncthreads();
            R fz[NR][NH];
            gather(pp, tag, fz);
            double acc[BicgA<NR>::NS];
#pragma unroll
            for (int k = 0; k < NR; ++k) acc[k] = 0.0;
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                if (!cs.active[k]) continue;
#pragma unroll
                for (int q = 0; q < RPT; ++q) {
                    if (!row[q]) continue;
                    const R y = product(k, q, pp, tag, fz);
                    v[k][q] = y;
                    acc[k] += (double)rh[k][q] * (double)y;
                }
            }
            tagged_reduce<BicgA<NR>, true>(a.tagged, &a.ctl->err, a.nwg, bid, gen, acc, false, red, s_err);
            ++gen;
            if (threadIdx.x == 0 && !s_err)
                for (int k = 0; k < NR; ++k)
                    if (sst[k].active) {
                        sst[k].alpha = sst[k].rho / red[k];
                        bc.coef[k] = (R)sst[k].alpha;
                    }
            __syncthreads();
            if (s_err) break;
            cs = bc;
        }
        // ---- phase B: s = r - alpha v; t = A s; x += alpha p; |s|^2, (t, s), (t, t) ----
        R s_[NR][RPT], t_[NR][RPT];
        {
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                const R al = cs.coef[k];
#pragma unroll
                for (int q = 0; q < RPT; ++q) {
                    s_[k][q] = R(0);
                    if (!cs.active[k] || !row[q]) continue;
                    const R sv = fma(-al, v[k][q], r[k][q]);
                    s_[k][q] = sv;
                    const int li = (int)threadIdx.x + q * kSolveThreads;
                    lv[k][li] = sv;
                    if (pub[q]) ps.put(0, k * n + lo + li, sv, tag);
                    x[k][q] += al * p[k][q];
                }
            }
            __syncthreads();
            R fz[NR][NH];
            gather(ps, tag, fz);
            double acc[BicgB<NR>::NS];
#pragma unroll
            for (int k = 0; k < BicgB<NR>::NS; ++k) acc[k] = 0.0;
#pragma unroll
            for (int k = 0; k < NR; ++k) {
#pragma unroll
                for (int q = 0; q < RPT; ++q) {
                    t_[k][q] = R(0);
                    if (!cs.active[k] || !row[q]) continue;
                    const R y = product(k, q, ps, tag, fz);
                    t_[k][q] = y;
                    const double sd = (double)s_[k][q], td = (double)y;
                    acc[3 * k] += sd * sd;
                    acc[3 * k + 1] += td * sd;
                    acc[3 * k + 2] += td * td;
                }
            }
            tagged_reduce<BicgB<NR>, true>(a.tagged, &a.ctl->err, a.nwg, bid, gen, acc, false, red, s_err);
            ++gen;
            if (threadIdx.x == 0 && !s_err)
                for (int k = 0; k < NR; ++k) {
                    RhsState &st = sst[k];
                    if (!st.active) continue;
                    const double nr = criterion(a.conv, red[3 * k], 0.0, 0.0, n);
                    st.used = it;
                    st.final_res = nr;
                    if (!isfinite(nr)) {
                        st.converged = 0; st.finite = 0; st.active = 0;
                    } else if (nr < a.tol) {
                        st.converged = 1; st.active = 0;
                    } else {
                        st.omega = red[3 * k + 1] / red[3 * k + 2];
                        bc.coef[k] = (R)st.omega;
                    }
                    bc.active[k] = st.active;
                }
            __syncthreads();
            if (s_err) break;
            cs = bc;
        }
        // ---- phase C: x += omega s; r = s - omega t; |r|^2, (rhat, r) ----
        {
            double acc[BicgC<NR>::NS];
#pragma unroll
            for (int k = 0; k < BicgC<NR>::NS; ++k) acc[k] = 0.0;
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                if (!cs.active[k]) continue;
                const R om = cs.coef[k];
#pragma unroll
                for (int q = 0; q < RPT; ++q) {
                    if (!row[q]) continue;
                    x[k][q] += om * s_[k][q];
                    const R rv = s_[k][q] - om * t_[k][q];
                    r[k][q] = rv;
                    acc[2 * k] += (double)rv * (double)rv;
                    acc[2 * k + 1] += (double)rh[k][q] * (double)rv;
                }
            }
            tagged_reduce<BicgC<NR>, true>(a.tagged, &a.ctl->err, a.nwg, bid, gen, acc, false, red, s_err);
            ++gen;
            if (threadIdx.x == 0 && !s_err)
                for (int k = 0; k < NR; ++k) {
                    RhsState &st = sst[k];
                    if (!st.active) continue;
                    const double nr = criterion(a.conv, red[2 * k], 0.0, 0.0, n);
                    st.final_res = nr;
                    if (nr < a.tol) {
                        st.converged = 1; st.used = it + 1; st.active = 0;
                    } else {
                        st.rho_prev = st.rho;
                        st.rho = red[2 * k + 1];
                        st.beta = (st.rho / st.rho_prev) * (st.alpha / st.omega);
                    }
                    bc.coef[k] = (R)st.beta;
                    bc.coef2[k] = (R)st.omega;
                    bc.active[k] = st.active;
                }
            __syncthreads();
            if (s_err) break;
            cs = bc;
        }
    }
    if (herr) st_coh(&a.ctl->err, 1);
#pragma unroll
    for (int q = 0; q < RPT; ++q)
        if (row[q])
#pragma unroll
            for (int k = 0; k < NR; ++k) a.x[(size_t)k * n + lo + (int)threadIdx.x + q * kSolveThreads] = x[k][q];
    if (bid == 0 && threadIdx.x < NR) a.st[threadIdx.x] = sst[threadIdx.x];
}

template <int RPT, int NZ, int NR>
__global__ __launch_bounds__(kSolveThreads) void k_bicg_local(SolveArgs<float> a) {
    bicg_local<RPT, NZ, NR>(a, blockIdx.x);
}

// batched launch (job lists as for k_cg_local_jobs)
template <int RPT, int NZ, int NR>
__global__ __launch_bounds__(kSolveThreads) void k_bicg_local_jobs(const SolveArgs<float> *__restrict__ args,
                                                                   const unsigned *__restrict__ job_off,
                                                                   const unsigned *__restrict__ jobs) {
    const unsigned j0 = job_off[blockIdx.x], j1 = job_off[blockIdx.x + 1];
    for (unsigned j = j0; j < j1; ++j) {
        const unsigned job = jobs[j];
        const SolveArgs<float> a = args[job >> 16];
        bicg_local<RPT, NZ, NR>(a, job & 0xffffu);
        __syncthreads();
    }
}

// ==========================================================================
// Mixed-precision iterative refinement (no counterpart in the reference, whose pressure
// solve is a single CG in the matrix dtype): r = b - A x in FP64 with x kept in FP64,
// the correction solved in the matrix dtype by the Krylov kernels above.
// ==========================================================================
constexpr int kRefineBlocks = 256;

// r = b - A x (FP64); part[4*blk + {0,1,2,3}] = sum r, sum r^2, sum b, sum b^2 of this workgroup's
// rows (grid-stride rows, fixed order: the host sums the parts in order)
template <typename R>
__global__ __launch_bounds__(256) void k_residual64(const int32_t *row, const int32_t *col, const R *val,
                                                    const R *b, const double *x, int n, double *r, double *part) {
    __shared__ double lds[4][4];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const double bi = (double)b[i];
        double s = bi;
        for (int e = row[i]; e < row[i + 1]; ++e) s -= (double)val[e] * x[col[e]];
        r[i] = s;
        acc[0] += s;
        acc[1] += s * s;
        acc[2] += bi;
        acc[3] += bi * bi;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double v = acc[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if ((threadIdx.x & 63) == 0) lds[k][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < 4) part[4 * blockIdx.x + threadIdx.x] =
        lds[threadIdx.x][0] + lds[threadIdx.x][1] + lds[threadIdx.x][2] + lds[threadIdx.x][3];
}

// largest |sum of a row| / max |entry of the row| over this workgroup's rows: ~0 when every row sums to
// zero, i.e. the constant vector spans a null space (pure Neumann or periodic pressure systems)
template <typename R>
__global__ __launch_bounds__(256) void k_row_sum_ratio(const int32_t *row, const R *val, int n, double *part) {
    __shared__ double lds[4];
    double m = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        double s = 0.0, a = 0.0;
        for (int e = row[i]; e < row[i + 1]; ++e) {
            s += (double)val[e];
            a = fmax(a, fabs((double)val[e]));
        }
        m = fmax(m, a > 0.0 ? fabs(s) / a : 1.0);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(lds[0], lds[1]), fmax(lds[2], lds[3]));
}

// out = scale * (r - shift) (rounded to the matrix dtype), d = 0
template <typename R>
__global__ __launch_bounds__(256) void k_refine_rhs(const double *r, double shift, double scale, R *out, R *d,
                                                   int n) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        out[i] = (R)((r[i] - shift) * scale);
        d[i] = R(0);
    }
}

// x += scale * d (FP64)
template <typename R>
__global__ __launch_bounds__(256) void k_refine_update(double *x, const R *d, double scale, int n) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) x[i] += scale * (double)d[i];
}

// a solve's final states and barrier error flag to the caller's slots (one launch after the solver kernel)
__global__ void k_solve_out(const RhsState *st, int nrhs, const int *err, RhsState *st_out, int *err_out) {
    const int k = threadIdx.x;
    if (k < nrhs) st_out[k] = st[k];
    if (k == 0) *err_out = *err;
}

// zero n 64-bit words (the tagged regions of k_cg_local, before its launch)
__global__ void k_zero64(u64 *p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}

// ==========================================================================
// BiCGStab without preconditioner (bicgstabSolveGPU, bicgstab_solver_kernel.cu:222-365)
// Three reduce-barriers per iteration: (rhat,v) | half-step norm, (t,s), (t,t) |
// norm, (rhat,r).  p = r + beta (p - omega v) and s = r - alpha v are folded
// into the following SpMVs, evaluated on the fly for the neighbours.  The vectors
// other workgroups read in the iterations (p, v, r) travel as agent-scope atomics,
// so the iteration barriers need no L2 writeback or invalidation.
// ==========================================================================
template <typename R, int NRHS>
__device__ __forceinline__ void bicgstab(const SolveArgs<R> &a, unsigned bid) {
    __shared__ double red[NRHS * kSlots];
    __shared__ RhsState sst[NRHS];
    __shared__ int s_err;
    const int n = a.n;
    const size_t stride = (size_t)a.nwg * kSolveThreads;
    const size_t i0 = (size_t)bid * kSolveThreads + threadIdx.x;
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
                // the row's entries in registers when they fit (one round trip), then each right-hand side's gathers
                RowRegs<R> rr;
                load_row(a, i, rr);
                const int rs = a.row[i], re = a.row[i + 1];
#pragma unroll
                for (int k = 0; k < NRHS; ++k) {
                    if (!act[k]) continue;
                    const size_t o = (size_t)k * n;
                    const R pn = fma(beta[k], fma(mom[k], ld_coh(vold + o + i), ld_coh(pold + o + i)), ld_coh(a.r + o + i));
                    st_coh(pnew + o + i, pn);
                    R s = 0;
                    if (rr.nz >= 0) {
                        R gv[kMaxNz], gp[kMaxNz], gr[kMaxNz];
#pragma unroll
                        for (int e = 0; e < kMaxNz; ++e) {
                            const size_t c = o + rr.col[e];
                            gv[e] = ld_coh(vold + c);
                            gp[e] = ld_coh(pold + c);
                            gr[e] = ld_coh(a.r + c);
                        }
#pragma unroll
                        for (int e = 0; e < kMaxNz; ++e)
                            if (e < rr.nz) s += rr.val[e] * fma(beta[k], fma(mom[k], gv[e], gp[e]), gr[e]);
                    } else {
                        for (int e = rs; e < re; ++e) {
                            const size_t c = o + a.col[e];
                            s += mat_val(a, e) * fma(beta[k], fma(mom[k], ld_coh(vold + c), ld_coh(pold + c)), ld_coh(a.r + c));
                        }
                    }
                    st_coh(vnew + o + i, s);
                    acc[k][0] += (double)a.rhat[o + i] * (double)s;
                }
            }
            PICT_FLAT_BARRIER_F(1, acc, IDS_D, MAX_D, false, {
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
                RowRegs<R> rr;
                load_row(a, i, rr);
                const int rs = a.row[i], re = a.row[i + 1];
#pragma unroll
                for (int k = 0; k < NRHS; ++k) {
                    if (!act[k]) continue;
                    const size_t o = (size_t)k * n;
                    const R mal = -alpha[k];
                    const R sv = fma(mal, ld_coh(vnew + o + i), ld_coh(a.r + o + i));
                    a.s[o + i] = sv;
                    R tv = 0;
                    if (rr.nz >= 0) {
                        R gv[kMaxNz], gr[kMaxNz];
#pragma unroll
                        for (int e = 0; e < kMaxNz; ++e) {
                            const size_t c = o + rr.col[e];
                            gv[e] = ld_coh(vnew + c);
                            gr[e] = ld_coh(a.r + c);
                        }
#pragma unroll
                        for (int e = 0; e < kMaxNz; ++e)
                            if (e < rr.nz) tv += rr.val[e] * fma(mal, gv[e], gr[e]);
                    } else {
                        for (int e = rs; e < re; ++e) {
                            const size_t c = o + a.col[e];
                            tv += mat_val(a, e) * fma(mal, ld_coh(vnew + c), ld_coh(a.r + c));
                        }
                    }
                    a.t[o + i] = tv;
                    a.x[o + i] += alpha[k] * ld_coh(pnew + o + i);
                    acc_res(acc, k, sv);
                    acc[k][3] += (double)tv * (double)sv;
                    acc[k][4] += (double)tv * (double)tv;
                }
            }
            PICT_FLAT_BARRIER_F(5, acc, IDS_B, MAX_B, false, {
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
                    st_coh(a.r + o + i, rv);
                    acc_res(acc, k, rv);
                    acc[k][3] += (double)a.rhat[o + i] * (double)rv;
                }
            }
            PICT_FLAT_BARRIER_F(4, acc, IDS_C, MAX_C, false, {
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
    if (bid == 0 && threadIdx.x == 0)
        for (int k = 0; k < NRHS; ++k) a.st[k] = sst[k];
}

template <typename R, int NRHS>
__global__ __launch_bounds__(kSolveThreads) void k_bicgstab(SolveArgs<R> a) {
    bicgstab<R, NRHS>(a, blockIdx.x);
}

// batched launch (job lists as for k_cg_local_jobs)
template <typename R, int NRHS>
__global__ __launch_bounds__(kSolveThreads) void k_bicgstab_jobs(const SolveArgs<R> *__restrict__ args,
                                                                 const unsigned *__restrict__ job_off,
                                                                 const unsigned *__restrict__ jobs) {
    const unsigned j0 = job_off[blockIdx.x], j1 = job_off[blockIdx.x + 1];
    for (unsigned j = j0; j < j1; ++j) {
        const unsigned job = jobs[j];
        const SolveArgs<R> a = args[job >> 16];
        bicgstab<R, NRHS>(a, job & 0xffffu);
        __syncthreads();
    }
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
thread_local bool g_force_textbook = false;  // pict_solve_refined's fallback for its correction solves
thread_local double g_last_ms = 0.0;
thread_local hipEvent_t g_ev[2] = {nullptr, nullptr};

// Occupancy (resident workgroups per CU) of a solver kernel on the current device, cached, at most
// kMaxResidentWgsPerCu: every workgroup of a persistent solver grid has to be resident at once.
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
    if (occ > kMaxResidentWgsPerCu) occ = kMaxResidentWgsPerCu;
    cache.push_back({dev, kernel, occ, cus});
    return PICT_OK;
}

enum class Kind { kCG1, kCGTextbook, kBiCG };

// stamp of a launch's foreign-read marks in k_cg_local (a stale mark equal to it only costs a store)
int next_salt() {
    static std::atomic<unsigned> launches{0};
    unsigned sv = ++launches;
    if (sv == 0) sv = ++launches;
    return (int)sv;
}

template <typename R, int NR>
const void *kernel_for(Kind kind) {
    switch (kind) {
        case Kind::kBiCG: return (const void *)k_bicgstab<R, NR>;
        case Kind::kCGTextbook: return (const void *)k_cg<R, NR>;
        default: return (const void *)k_cg1<R, NR, false>;
    }
}
template <typename R>
const void *cg1_lean_for_nrhs(int nrhs) {
    switch (nrhs) {
        case 1: return (const void *)k_cg1<R, 1, true>;
        case 2: return (const void *)k_cg1<R, 2, true>;
        case 3: return (const void *)k_cg1<R, 3, true>;
        default: return (const void *)k_cg1<R, 4, true>;
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
// k_cg_local instance for rows per thread `rpt` and cached row length `nz`
template <typename R, int NZ, bool FULL>
const void *local_cg_kernel_nz(int rpt) {
    switch (rpt) {
        case 1: return (const void *)k_cg_local<R, 1, NZ, FULL>;
        case 2: return (const void *)k_cg_local<R, 2, NZ, FULL>;
        case 4: return (const void *)k_cg_local<R, 4, NZ, FULL>;
        default:
            if constexpr (sizeof(R) == 4 && NZ <= 5) return (const void *)k_cg_local<R, 8, NZ, FULL>;
            return nullptr;
    }
}
// `full`: the convergence criterion needs the sum and maximum of |r|
template <typename R>
const void *local_cg_kernel(int rpt, int nz, bool full) {
    if (full) return nz <= 5 ? local_cg_kernel_nz<R, 5, true>(rpt) : local_cg_kernel_nz<R, 7, true>(rpt);
    return nz <= 5 ? local_cg_kernel_nz<R, 5, false>(rpt) : local_cg_kernel_nz<R, 7, false>(rpt);
}

// k_bicg_local instances: 5-entry rows with one or two right-hand sides, 7-entry rows with one
template <int NZ, int NR>
const void *bicg_local_kernel_nz(int rpt) {
    switch (rpt) {
        case 1: return (const void *)k_bicg_local<1, NZ, NR>;
        case 2: return (const void *)k_bicg_local<2, NZ, NR>;
        case 4: return (const void *)k_bicg_local<4, NZ, NR>;
        default: return nullptr;
    }
}
const void *bicg_local_kernel(int rpt, int nz, int nrhs) {
    if (nz <= 5) return nrhs == 1 ? bicg_local_kernel_nz<5, 1>(rpt) : nrhs == 2 ? bicg_local_kernel_nz<5, 2>(rpt) : nullptr;
    return nrhs == 1 ? bicg_local_kernel_nz<7, 1>(rpt) : nullptr;
}
// Solver grids of at most one workgroup per compute unit (the device's queried CU count; every solver kernel has
// an occupancy of at least one) are launched as ordinary kernels: the cooperative launch costs host time that
// leaves the GPU idle between a solve and the next kernel (periodic box 256x256 forward+backward: 18.3 M against
// 16.8 M cells*steps/s).  Residency: on an idle device the dispatcher places all of a grid of at most one
// workgroup per CU at once.  Work of other queues that holds CUs delays the workgroups that find no room, and
// they start when it releases them; kernels that never wait on other grids (every stage kernel, torch's kernels)
// always do, so the grid becomes fully resident and its barriers complete.  Two persistent grids that wait on their own barriers and
// together exceed what the CUs hold (two such solves on concurrent streams, or in two processes sharing the GPU)
// can each be partly resident; the spin limit then ends the waiting solve with PICT_ERR_DEVICE_SYNC instead of a
// hang.  The library's own solves are queued on one stream per call, so this only arises from concurrent callers.
// PICT_SOLVER_PLAIN_LAUNCH=0 keeps the cooperative launch, whose residency the runtime checks.
static bool plain_launch_enabled() {
    const char *e = getenv("PICT_SOLVER_PLAIN_LAUNCH");
    return !e || atoi(e) != 0;
}
// PICT_BICG_LOCAL=0 keeps BiCGStab on the streaming kernel
static bool bicg_local_enabled() {
    const char *e = getenv("PICT_BICG_LOCAL");
    return !e || atoi(e) != 0;
}

// Workgroup budget of the register-resident CG: the cost of its barrier grows with the
// number of workgroups, the per-thread work with rows per thread.
// (PICT_CG_LOCAL_WGS: budget; 0 turns the kernel off)
static int local_cg_target_wgs() {
    const char *e = getenv("PICT_CG_LOCAL_WGS");
    return e ? atoi(e) : 64;
}

// Which kernel a solve runs, and on how many workgroups.
struct Plan {
    Kind kind;
    const void *kernel;
    long long nwg;
    bool local;       // k_cg_local
    bool bicg_local;  // k_bicg_local
    int rpt, nz;      // register-resident kernels: rows per thread, cached row length
    bool full;      // k_cg_local reduces the sum and maximum of |r| too
    bool lean;      // k_cg1 without the register row cache
    int cus;        // compute units of the device
};

template <typename R>
pict_status plan_solve(const pict_csr_desc *A, int nrhs, const pict_solve_opts *o, Plan &pl) {
    const int n = A->n;
    // PICT_CG_TEXTBOOK=1 selects the two-barrier textbook CG kernel (diagnostics / parity checks)
    const char *tb_env = getenv("PICT_CG_TEXTBOOK");
    const bool textbook_cg = (tb_env && atoi(tb_env) != 0) || g_force_textbook;
    pl = Plan{};
    // CG with residual resets runs the textbook kernel: the reference resets the residual only in the driver's
    // non-orthogonal branch (PISOtorch_simulation.py:1280), whose pressure matrix is not symmetric, and on a
    // non-symmetric matrix the single-reduction recurrences no longer follow textbook CG's iterates (they stall
    // where textbook CG, the reference's algorithm, converges)
    pl.kind = o->use_bicg ? Kind::kBiCG
              : (textbook_cg || o->residual_reset_steps > 0) ? Kind::kCGTextbook
                                                             : Kind::kCG1;
    pl.kernel = kernel_for_nrhs<R>(pl.kind, nrhs);
    int occ = 0, cus = 0;
    pict_status st = occupancy(pl.kernel, occ, cus);
    if (st) return st;
    pl.cus = cus;
    if (occ < 1) {
        set_error("solver kernel cannot be resident (occupancy 0)");
        return PICT_ERR_HIP;
    }
    // one right-hand side of the single-reduction CG without residual resets: the
    // register-resident kernel when its workgroups fit on the device at once
    const int target = local_cg_target_wgs();
    if (pl.kind == Kind::kCG1 && nrhs == 1 && o->residual_reset_steps <= 0 && target > 0) {
        const int nz = (long long)A->nnz <= 5LL * n ? 5 : 7;
        const int rmax = sizeof(R) == 4 && nz <= 5 ? 8 : 4;  // register budget (FP64 rows, 7-entry rows)
        const bool full = o->convergence_criterion != PICT_CONV_NORM2 &&
                          o->convergence_criterion != PICT_CONV_NORM2_NORMALIZED;
        for (int rpt = 1; rpt <= rmax; rpt *= 2) {
            const void *k = local_cg_kernel<R>(rpt, nz, full);
            const long long per = (long long)kSolveThreads * rpt;
            const long long wgs = ((long long)n + per - 1) / per;
            if (wgs > target && rpt < rmax) continue;
            int locc = 0, lcus = 0;
            st = occupancy(k, locc, lcus);
            if (st) return st;
            if (wgs <= (long long)locc * lcus) {
                pl.kernel = k;
                pl.nwg = wgs;
                pl.local = true;
                pl.rpt = rpt;
                pl.nz = nz;
                pl.full = full;
                return PICT_OK;
            }
            break;
        }
    }
    // BiCGStab with FP32 rows in registers (norm-2 criteria, up to two right-hand sides), on the smallest rows per
    // thread that keep the grid within the workgroup budget and resident
    if constexpr (sizeof(R) == 4) {
        const bool norm2 = o->convergence_criterion == PICT_CONV_NORM2 ||
                           o->convergence_criterion == PICT_CONV_NORM2_NORMALIZED;
        if (pl.kind == Kind::kBiCG && norm2 && target > 0 && bicg_local_enabled()) {
            const int nz = (long long)A->nnz <= 5LL * n ? 5 : 7;
            for (int rpt = 1; rpt <= 4; rpt *= 2) {
                const void *k = bicg_local_kernel(rpt, nz, nrhs);
                if (!k) break;
                const long long per = (long long)kSolveThreads * rpt;
                const long long wgs = ((long long)n + per - 1) / per;
                if (wgs > target && rpt < 4) continue;
                int locc = 0, lcus = 0;
                st = occupancy(k, locc, lcus);
                if (st) return st;
                if (wgs <= (long long)locc * lcus) {
                    pl.kernel = k;
                    pl.nwg = wgs;
                    pl.bicg_local = true;
                    pl.rpt = rpt;
                    pl.nz = nz;
                    return PICT_OK;
                }
                break;
            }
        }
    }
    // rows per thread: more rows per workgroup means fewer barrier participants
    static const int rows_per_thread = [] {
        const char *e = getenv("PICT_SOLVER_ROWS_PER_THREAD");
        const int v = e ? atoi(e) : 2;
        return v < 1 ? 1 : v;
    }();
    const long long per_wg = (long long)kSolveThreads * rows_per_thread;
    auto grid = [&](int occ_k) {
        long long nwg = ((long long)n + per_wg - 1) / per_wg;
        const long long cap = (long long)occ_k * cus;
        if (nwg > cap) nwg = cap;
        if (nwg > 2048) nwg = 2048;
        if (nwg < 1) nwg = 1;
        return nwg;
    };
    pl.nwg = grid(occ);
    // rows beyond the register row cache: the lean kernel, with the waves its registers allow
    if (pl.kind == Kind::kCG1 && (long long)pl.nwg * kSolveThreads * kRowsCached < (long long)n) {
        const void *k = cg1_lean_for_nrhs<R>(nrhs);
        int locc = 0, lcus = 0;
        st = occupancy(k, locc, lcus);
        if (st) return st;
        if (locc > occ) {
            pl.kernel = k;
            pl.lean = true;
            pl.nwg = grid(locc);
            // a multiple of eight workgroups: the lean kernel places the chunks of one XCD next to each other
            if (pl.nwg >= 16) pl.nwg -= pl.nwg % 8;
        }
    }
    return PICT_OK;
}

// Workspace of one solve: vectors | FP64 partials | final states | transpose map | control block |
// k_cg_local's tagged partials and published rows.  The control block and the regions after it
// are zeroed by one fill kernel before the launch.
struct Layout {
    size_t vec, off_part, off_st, off_perm, off_ctl, off_zero, words_tagged, words_zero, total;
};

template <typename R>
Layout layout_of(const pict_csr_desc *A, int nrhs, const pict_solve_opts *o, const Plan &pl) {
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    Layout L;
    const size_t n = (size_t)A->n;
    L.vec = n * nrhs;
    const int nvec = pl.kind == Kind::kCGTextbook ? (o->return_best_result ? 5 : 4) : 9;
    const size_t bytes_part = sizeof(double) * partial_doubles(pl.nwg);
    const size_t bytes_perm = (o->transpose_a && !A->transpose_perm) ? sizeof(int32_t) * (size_t)A->nnz : 0;
    L.off_part = al(sizeof(R) * L.vec * nvec);
    L.off_st = L.off_part + al(bytes_part);
    L.off_perm = L.off_st + al(sizeof(RhsState) * nrhs);
    L.off_ctl = L.off_perm + al(bytes_perm);
    L.off_zero = L.off_ctl + al(sizeof(Ctl));
    L.words_tagged = pl.local || pl.bicg_local ? tagged_words(pl.nwg) : 0;
    L.words_zero = pl.local        ? L.words_tagged + 6 * n
                   : pl.bicg_local ? L.words_tagged + 2 * (size_t)nrhs * n
                                   : 0;
    L.total = L.off_zero + sizeof(unsigned long long) * L.words_zero + 256;
    return L;
}

template <typename R>
pict_status enqueue_typed(const pict_csr_desc *A, const void *rhs, void *x, int nrhs, const pict_solve_opts *o,
                          RhsState *state_out, int *err_out, hipEvent_t ev0, hipEvent_t ev1, void *ws_in,
                          size_t ws_bytes, hipStream_t stream) {
    const int n = A->n;
    Plan pl;
    pict_status st = plan_solve<R>(A, nrhs, o, pl);
    if (st) return st;
    const Kind kind = pl.kind;
    const void *kernel = pl.kernel;
    const long long nwg = pl.nwg;
    const Layout L = layout_of<R>(A, nrhs, o, pl);
    const size_t vec = L.vec;

    char *ws = (char *)ws_in;
    struct Guard {
        char *p; hipStream_t s;
        ~Guard() { if (p) (void)ws_free(p, s); }
    } guard{nullptr, stream};
    if (ws) {
        PICT_CHECK_ARG(ws_bytes >= L.total, "solver workspace too small (%zu < %zu bytes)", ws_bytes, L.total);
    } else {
        PICT_HIP_TRY(ws_alloc((void **)&ws, L.total, stream));
        guard.p = ws;
    }
    const size_t off_part = L.off_part, off_ctl = L.off_ctl, off_st = L.off_st, off_perm = L.off_perm,
                 off_zero = L.off_zero, words_tagged = L.words_tagged;

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
    a.crit_scale = o->convergence_criterion == PICT_CONV_NORM2_NORMALIZED ? 1.0 / sqrt((double)n)
                   : o->convergence_criterion == PICT_CONV_ABS_MEAN      ? 1.0 / (double)n
                                                                         : 1.0;
    a.tagged = reinterpret_cast<unsigned long long *>(ws + off_zero);
    a.pub = a.tagged + words_tagged;
    a.salt = next_salt();
    {
        // control block (+ k_cg_local's tagged regions): a fill kernel, which a captured graph replays safely
        const size_t words = (off_zero - off_ctl) / 8 + L.words_zero;
        const unsigned blocks = (unsigned)std::min<size_t>((words + 255) / 256, 1024);
        hipLaunchKernelGGL(k_zero64, dim3(blocks), dim3(256), 0, stream, (unsigned long long *)(ws + off_ctl), words);
        PICT_HIP_TRY(hipGetLastError());
    }
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
    if (plain_launch_enabled() && nwg <= pl.cus)
        PICT_HIP_TRY(hipLaunchKernel(kernel, dim3((unsigned)nwg), dim3(kSolveThreads), args, 0, stream));
    else
        PICT_HIP_TRY(hipLaunchCooperativeKernel(kernel, dim3((unsigned)nwg), dim3(kSolveThreads), args, 0, stream));
    if (ev1) PICT_HIP_TRY(hipEventRecord(ev1, stream));
    hipLaunchKernelGGL(k_solve_out, dim3(1), dim3(64), 0, stream, (const RhsState *)a.st, nrhs,
                       (const int *)&a.ctl->err, state_out, err_out);
    PICT_HIP_TRY(hipGetLastError());
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
    return PICT_OK;
}

// The ILU(0)-preconditioned BiCGStab runs only through the synchronous entry points (its analysis reads the
// pattern on the host); the enqueue-only paths refuse it.
pict_status refuse_preconditioned(const pict_solve_opts *o) {
    if (o->use_bicg && o->bicg_with_preconditioner) {
        set_error("the ILU(0)-preconditioned BiCGStab runs through pict_solve_linear only");
        return PICT_ERR_UNSUPPORTED;
    }
    return PICT_OK;
}

}  // namespace

namespace pict {

size_t solve_state_bytes() { return sizeof(RhsState); }

pict_status solve_workspace_bytes(const pict_csr_desc *A, int nrhs, const pict_solve_opts *o, size_t *bytes) {
    pict_status st = validate_solve(A, nrhs, o);
    if (st || (st = refuse_preconditioned(o))) return st;
    Plan pl;
    if (A->dtype == PICT_DTYPE_F32) {
        if ((st = plan_solve<float>(A, nrhs, o, pl))) return st;
        *bytes = layout_of<float>(A, nrhs, o, pl).total;
    } else {
        if ((st = plan_solve<double>(A, nrhs, o, pl))) return st;
        *bytes = layout_of<double>(A, nrhs, o, pl).total;
    }
    return PICT_OK;
}

pict_status solve_enqueue(const pict_csr_desc *A, const void *rhs, void *x, int nrhs, const pict_solve_opts *o,
                          void *state_out, int *err_out, hipEvent_t ev0, hipEvent_t ev1, hipStream_t stream,
                          void *ws, size_t ws_bytes) {
    pict_status st = validate_solve(A, nrhs, o);
    if (st || (st = refuse_preconditioned(o))) return st;
    if (A->dtype == PICT_DTYPE_F32)
        return enqueue_typed<float>(A, rhs, x, nrhs, o, (RhsState *)state_out, err_out, ev0, ev1, ws, ws_bytes,
                                    stream);
    return enqueue_typed<double>(A, rhs, x, nrhs, o, (RhsState *)state_out, err_out, ev0, ev1, ws, ws_bytes,
                                 stream);
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

// ---------------------------------------------------------------------------
// Batched solves: `count` independent systems (same dtype, right-hand sides per
// system and options, any sizes) in one cooperative launch of a job-list kernel.
// Each system runs the single-system algorithm on its own workspace; the host fixes
// every system's workgroup count and deals the (system, workgroup) jobs to the
// launch's resident workgroups by list scheduling: systems in order of decreasing
// estimated cost, each on the workgroups that become free earliest.
// ---------------------------------------------------------------------------
struct BatchSolve {
    int count = 0, nrhs = 0, dtype = 0;
    bool local = false;       // k_cg_local_jobs
    bool bicg_local = false;  // k_bicg_local_jobs (else, without `local`, k_bicgstab_jobs)
    int rpt = 0, nz = 0, rows_per_wg = 0;
    const void *kernel = nullptr;
    unsigned grid = 0;   // workgroups of the launch, all resident
    std::vector<int> n, nwg;
    size_t off_args = 0, off_joboff = 0, off_jobs = 0, off_states = 0, off_err = 0, off_zero = 0;
    size_t zero_words = 0, total = 0, args_stride = 0;
    char *arena = nullptr;
    std::vector<char> args_host;
    std::vector<unsigned> job_off, jobs;
    double makespan = 0.0;
    ~BatchSolve() {
        if (arena) (void)hipFree(arena);
    }
};

namespace {

template <typename R, int NZ>
const void *local_cg_jobs_kernel_nz(int rpt) {
    switch (rpt) {
        case 1: return (const void *)k_cg_local_jobs<R, 1, NZ, false>;
        case 2: return (const void *)k_cg_local_jobs<R, 2, NZ, false>;
        case 4: return (const void *)k_cg_local_jobs<R, 4, NZ, false>;
        case 8:
            if constexpr (sizeof(R) == 4 && NZ <= 5) return (const void *)k_cg_local_jobs<R, 8, NZ, false>;
            return nullptr;
        default: return nullptr;
    }
}
template <int NZ, int NR>
const void *bicg_local_jobs_kernel_nz(int rpt) {
    switch (rpt) {
        case 1: return (const void *)k_bicg_local_jobs<1, NZ, NR>;
        case 2: return (const void *)k_bicg_local_jobs<2, NZ, NR>;
        case 4: return (const void *)k_bicg_local_jobs<4, NZ, NR>;
        default: return nullptr;
    }
}
// register-resident BiCGStab instances: FP32, 5-entry rows with one or two right-hand sides, 7-entry rows
// with one
const void *bicg_local_jobs_kernel(int rpt, int nz, int nrhs) {
    if (nz <= 5) return nrhs == 1 ? bicg_local_jobs_kernel_nz<5, 1>(rpt) : nrhs == 2 ? bicg_local_jobs_kernel_nz<5, 2>(rpt) : nullptr;
    return nrhs == 1 ? bicg_local_jobs_kernel_nz<7, 1>(rpt) : nullptr;
}

template <typename R>
const void *bicgstab_jobs_kernel(int nrhs) {
    switch (nrhs) {
        case 1: return (const void *)k_bicgstab_jobs<R, 1>;
        case 2: return (const void *)k_bicgstab_jobs<R, 2>;
        case 3: return (const void *)k_bicgstab_jobs<R, 3>;
        default: return (const void *)k_bicgstab_jobs<R, 4>;
    }
}

// OR of the systems' barrier error flags
template <typename R>
__global__ __launch_bounds__(256) void k_batch_err(const SolveArgs<R> *args, int count, int *out) {
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    for (int s = threadIdx.x; s < count; s += blockDim.x)
        if (__hip_atomic_load(&args[s].ctl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicOr(&any, 1);
    __syncthreads();
    if (threadIdx.x == 0) *out = any;
}

// List scheduling of the systems' jobs onto b.grid workgroups (cost: estimated time of each solve).
void schedule(BatchSolve &b, const std::vector<double> &cost) {
    const unsigned G = b.grid;
    std::vector<int> order(b.count);
    for (int s = 0; s < b.count; ++s) order[s] = s;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
    std::vector<double> avail(G, 0.0);
    std::vector<std::vector<unsigned>> lists(G);
    std::vector<unsigned> idx(G);
    b.makespan = 0.0;
    for (int s : order) {
        const unsigned k = (unsigned)b.nwg[s];
        for (unsigned g = 0; g < G; ++g) idx[g] = g;
        auto earlier = [&](unsigned x, unsigned y) { return avail[x] < avail[y] || (avail[x] == avail[y] && x < y); };
        std::nth_element(idx.begin(), idx.begin() + (k - 1), idx.end(), earlier);
        std::sort(idx.begin(), idx.begin() + k);
        double start = 0.0;
        for (unsigned j = 0; j < k; ++j) start = std::max(start, avail[idx[j]]);
        const double fin = start + cost[s];
        for (unsigned j = 0; j < k; ++j) {
            lists[idx[j]].push_back(((unsigned)s << 16) | j);
            avail[idx[j]] = fin;
        }
        b.makespan = std::max(b.makespan, fin);
    }
    b.job_off.assign(G + 1, 0);
    b.jobs.clear();
    for (unsigned g = 0; g < G; ++g) {
        b.job_off[g] = (unsigned)b.jobs.size();
        b.jobs.insert(b.jobs.end(), lists[g].begin(), lists[g].end());
    }
    b.job_off[G] = (unsigned)b.jobs.size();
}

// default schedule estimate: CG iterations grow like the square root of the size (2-D and 3-D
// Poisson-like systems); a BiCGStab iteration streams its rows, so its time grows with the rows
// per workgroup on top of the barrier latency
std::vector<double> default_cost(const BatchSolve &b) {
    std::vector<double> c(b.count);
    for (int s = 0; s < b.count; ++s)
        c[s] = b.local || b.bicg_local ? std::sqrt((double)b.n[s]) : (2048.0 + (double)b.n[s] / b.nwg[s]);
    return c;
}

template <typename R>
void fill_args(BatchSolve &b, int s, const pict_csr_desc &A, const void *rhs, void *x, const pict_solve_opts *o,
               const std::vector<size_t> &off_vec, const std::vector<size_t> &off_part,
               const std::vector<size_t> &off_ctl) {
    auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
    SolveArgs<R> a;
    memset(&a, 0, sizeof a);
    const size_t vec = (size_t)A.n * b.nrhs;
    a.row = A.row;
    a.col = A.index;
    a.val = (const R *)A.value;
    a.b = (const R *)rhs;
    a.x = (R *)x;
    R *v = (R *)(b.arena + off_vec[s]);
    R **vs[9] = {&a.r, &a.p0, &a.p1, &a.v0, &a.rhat, &a.v1, &a.s, &a.t, &a.best};
    for (R **pp : vs) {
        *pp = v;
        v += vec;
    }
    a.partials = (double *)(b.arena + off_part[s]);
    a.ctl = (Ctl *)(b.arena + off_ctl[s]);
    a.st = (RhsState *)(b.arena + b.off_states) + (size_t)s * b.nrhs;
    a.n = A.n;
    a.nrhs = b.nrhs;
    a.maxit = o->max_iterations;
    a.conv = o->convergence_criterion;
    a.reset_steps = o->residual_reset_steps;
    a.return_best = o->return_best_result;
    a.zero_shortcut = o->zero_rhs_shortcut;
    a.tol = o->tolerance;
    a.nwg = (unsigned)b.nwg[s];
    a.crit_scale = o->convergence_criterion == PICT_CONV_NORM2_NORMALIZED ? 1.0 / sqrt((double)A.n)
                   : o->convergence_criterion == PICT_CONV_ABS_MEAN      ? 1.0 / (double)A.n
                                                                         : 1.0;
    a.tagged = reinterpret_cast<unsigned long long *>(b.arena + off_ctl[s] + al(sizeof(Ctl)));
    a.pub = a.tagged + (b.local || b.bicg_local ? tagged_words(b.nwg[s]) : 0);
    a.salt = next_salt();
    memcpy(b.args_host.data() + (size_t)s * sizeof(SolveArgs<R>), &a, sizeof a);
}

template <typename R>
pict_status batch_create_typed(const pict_csr_desc *A, const void *const *rhs, void *const *x, int count, int nrhs,
                               const pict_solve_opts *o, const BatchSolveConfig &cfg, const double *cost,
                               BatchSolve **out, hipStream_t stream) {
    auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
    std::unique_ptr<BatchSolve> b(new BatchSolve());
    b->count = count;
    b->nrhs = nrhs;
    b->dtype = A[0].dtype;
    b->local = !o->use_bicg;
    int occ = 0, cus = 0;
    if (b->local) {
        PICT_CHECK_ARG(nrhs == 1, "batched CG solves take one right-hand side per system");
        PICT_CHECK_ARG(o->residual_reset_steps <= 0, "batched CG solves do not reset the residual");
        if (o->convergence_criterion != PICT_CONV_NORM2 && o->convergence_criterion != PICT_CONV_NORM2_NORMALIZED) {
            set_error("batched CG solves support the norm-2 convergence criteria");
            return PICT_ERR_UNSUPPORTED;
        }
        bool all5 = true;
        for (int s = 0; s < count; ++s) all5 = all5 && (long long)A[s].nnz <= 5LL * A[s].n;
        b->nz = all5 ? 5 : 7;
        const int rmax = sizeof(R) == 4 && b->nz <= 5 ? 8 : 4;
        b->rpt = cfg.rows_per_thread > 0 ? cfg.rows_per_thread : (rmax >= 4 ? 4 : rmax);
        PICT_CHECK_ARG(b->rpt == 1 || b->rpt == 2 || b->rpt == 4 || b->rpt == 8, "rows per thread must be 1, 2, 4 or 8");
        PICT_CHECK_ARG(b->rpt <= rmax, "rows per thread %d exceeds the register budget (%d)", b->rpt, rmax);
        b->kernel = b->nz <= 5 ? local_cg_jobs_kernel_nz<R, 5>(b->rpt) : local_cg_jobs_kernel_nz<R, 7>(b->rpt);
    } else {
        bool all5 = true;
        for (int s = 0; s < count; ++s) all5 = all5 && (long long)A[s].nnz <= 5LL * A[s].n;
        b->nz = all5 ? 5 : 7;
        const bool norm2 = o->convergence_criterion == PICT_CONV_NORM2 ||
                           o->convergence_criterion == PICT_CONV_NORM2_NORMALIZED;
        if (cfg.bicg_rows_per_thread > 0 && sizeof(R) == 4 && norm2 &&
            bicg_local_jobs_kernel(cfg.bicg_rows_per_thread, b->nz, nrhs)) {
            b->bicg_local = true;
            b->rpt = cfg.bicg_rows_per_thread;
            b->kernel = bicg_local_jobs_kernel(b->rpt, b->nz, nrhs);
        } else {
            b->rows_per_wg = cfg.rows_per_wg > 0 ? cfg.rows_per_wg : 1024;
            PICT_CHECK_ARG(b->rows_per_wg >= 64, "rows per workgroup must be at least 64");
            b->kernel = bicgstab_jobs_kernel<R>(nrhs);
        }
    }
    PICT_CHECK_ARG(b->kernel != nullptr, "no batched solver kernel for this configuration");
    pict_status st = occupancy(b->kernel, occ, cus);
    if (st) return st;
    if (occ < 1) {
        set_error("batched solver kernel cannot be resident (occupancy 0)");
        return PICT_ERR_HIP;
    }
    const long long cap = (long long)occ * cus;
    long long sum = 0;
    b->n.resize(count);
    b->nwg.resize(count);
    for (int s = 0; s < count; ++s) {
        const long long n = A[s].n;
        const bool resident = b->local || b->bicg_local;
        long long w = resident ? (n + (long long)kSolveThreads * b->rpt - 1) / ((long long)kSolveThreads * b->rpt)
                               : (n + b->rows_per_wg - 1) / b->rows_per_wg;
        if (w < 1) w = 1;
        if (!resident && w > cap) w = cap;
        if (w > cap || w > 65535) {
            set_error("system %d (%lld rows) needs %lld workgroups; at most %lld are resident", s, n, w, cap);
            return PICT_ERR_UNSUPPORTED;
        }
        b->n[s] = (int)n;
        b->nwg[s] = (int)w;
        sum += w;
    }
    b->grid = (unsigned)std::min(sum, cap);
    // arena: tables | states | error flag | per-system vectors and partials | per-system control blocks and
    // tagged regions (contiguous: one fill kernel clears them before each launch)
    b->args_stride = sizeof(SolveArgs<R>);
    size_t off = 0;
    b->off_args = off;
    off += al(b->args_stride * count);
    b->off_joboff = off;
    off += al(sizeof(unsigned) * (b->grid + 1));
    b->off_jobs = off;
    off += al(sizeof(unsigned) * (size_t)sum);
    b->off_states = off;
    off += al(sizeof(RhsState) * (size_t)count * nrhs);
    b->off_err = off;
    off += 256;
    std::vector<size_t> off_vec(count), off_part(count), off_ctl(count);
    for (int s = 0; s < count; ++s) {
        off_vec[s] = off;
        off += al(sizeof(R) * (size_t)b->n[s] * nrhs * 9);
        off_part[s] = off;
        off += al(sizeof(double) * partial_doubles(b->nwg[s]));
    }
    b->off_zero = off;
    for (int s = 0; s < count; ++s) {
        off_ctl[s] = off;
        const size_t words = b->local        ? tagged_words(b->nwg[s]) + 6 * (size_t)b->n[s]
                             : b->bicg_local ? tagged_words(b->nwg[s]) + 2 * (size_t)nrhs * b->n[s]
                                             : 0;
        off += al(sizeof(Ctl)) + al(sizeof(unsigned long long) * words);
    }
    b->zero_words = (off - b->off_zero) / sizeof(unsigned long long);
    b->total = off;
    PICT_HIP_TRY(hipMalloc((void **)&b->arena, b->total));
    b->args_host.assign(b->args_stride * count, 0);
    for (int s = 0; s < count; ++s) fill_args<R>(*b, s, A[s], rhs[s], x[s], o, off_vec, off_part, off_ctl);
    std::vector<double> c = cost ? std::vector<double>(cost, cost + count) : default_cost(*b);
    schedule(*b, c);
    PICT_HIP_TRY(hipMemcpyAsync(b->arena + b->off_args, b->args_host.data(), b->args_host.size(),
                                hipMemcpyHostToDevice, stream));
    PICT_HIP_TRY(hipMemcpyAsync(b->arena + b->off_joboff, b->job_off.data(), sizeof(unsigned) * b->job_off.size(),
                                hipMemcpyHostToDevice, stream));
    PICT_HIP_TRY(hipMemcpyAsync(b->arena + b->off_jobs, b->jobs.data(), sizeof(unsigned) * b->jobs.size(),
                                hipMemcpyHostToDevice, stream));
    PICT_HIP_TRY(hipStreamSynchronize(stream));
    *out = b.release();
    return PICT_OK;
}

}  // namespace

pict_status batch_solve_create(const pict_csr_desc *A, const void *const *rhs, void *const *x, int count, int nrhs,
                               const pict_solve_opts *o, const BatchSolveConfig &cfg, const double *cost,
                               BatchSolve **out, hipStream_t stream) {
    PICT_CHECK_ARG(A && rhs && x && o && out, "NULL argument");
    PICT_CHECK_ARG(count >= 1 && count <= 65535, "batch size must be 1..65535 (got %d)", count);
    for (int s = 0; s < count; ++s) {
        pict_status st = validate_solve(&A[s], nrhs, o);
        if (st) return st;
        PICT_CHECK_ARG(A[s].dtype == A[0].dtype, "system %d: dtype differs from system 0", s);
        PICT_CHECK_ARG(rhs[s] && x[s], "system %d: vectors missing", s);
    }
    PICT_CHECK_ARG(!o->transpose_a, "batched solves are not transposed");
    return A[0].dtype == PICT_DTYPE_F32 ? batch_create_typed<float>(A, rhs, x, count, nrhs, o, cfg, cost, out, stream)
                                        : batch_create_typed<double>(A, rhs, x, count, nrhs, o, cfg, cost, out, stream);
}

void batch_solve_destroy(BatchSolve *b) { delete b; }

pict_status batch_solve_update(BatchSolve *b, const pict_csr_desc *A, const void *const *rhs, void *const *x,
                               hipStream_t stream) {
    bool changed = false;
    for (int s = 0; s < b->count; ++s) {
        char *h = b->args_host.data() + (size_t)s * b->args_stride;
        // the leading members of SolveArgs: row, col, perm, val, b, x
        const void *want[6] = {A[s].row, A[s].index, nullptr, A[s].value, rhs[s], x[s]};
        PICT_CHECK_ARG(A[s].n == b->n[s], "system %d changed size", s);
        if (memcmp(h, want, sizeof want) != 0) {
            memcpy(h, want, sizeof want);
            changed = true;
        }
    }
    if (!changed) return PICT_OK;
    // queued launches read the table: replace it only when they are done
    PICT_HIP_TRY(hipStreamSynchronize(stream));
    PICT_HIP_TRY(hipMemcpyAsync(b->arena + b->off_args, b->args_host.data(), b->args_host.size(),
                                hipMemcpyHostToDevice, stream));
    PICT_HIP_TRY(hipStreamSynchronize(stream));
    return PICT_OK;
}

pict_status batch_solve_reschedule(BatchSolve *b, const double *cost, hipStream_t stream) {
    std::vector<double> c(cost, cost + b->count);
    schedule(*b, c);
    PICT_HIP_TRY(hipStreamSynchronize(stream));
    PICT_HIP_TRY(hipMemcpyAsync(b->arena + b->off_joboff, b->job_off.data(), sizeof(unsigned) * b->job_off.size(),
                                hipMemcpyHostToDevice, stream));
    PICT_HIP_TRY(hipMemcpyAsync(b->arena + b->off_jobs, b->jobs.data(), sizeof(unsigned) * b->jobs.size(),
                                hipMemcpyHostToDevice, stream));
    PICT_HIP_TRY(hipStreamSynchronize(stream));
    return PICT_OK;
}

pict_status batch_solve_enqueue(const BatchSolve *b, void *states_out, int *err_out, hipEvent_t ev0, hipEvent_t ev1,
                                hipStream_t stream) {
    {
        const unsigned blocks = (unsigned)std::min<size_t>((b->zero_words + 255) / 256, 2048);
        hipLaunchKernelGGL(k_zero64, dim3(blocks ? blocks : 1), dim3(256), 0, stream,
                           (unsigned long long *)(b->arena + b->off_zero), b->zero_words);
        PICT_HIP_TRY(hipGetLastError());
    }
    void *args_dev = b->arena + b->off_args;
    unsigned *job_off = (unsigned *)(b->arena + b->off_joboff), *jobs = (unsigned *)(b->arena + b->off_jobs);
    void *kargs[] = {&args_dev, &job_off, &jobs};
    if (ev0) PICT_HIP_TRY(hipEventRecord(ev0, stream));
    PICT_HIP_TRY(hipLaunchCooperativeKernel(b->kernel, dim3(b->grid), dim3(kSolveThreads), kargs, 0, stream));
    if (ev1) PICT_HIP_TRY(hipEventRecord(ev1, stream));
    int *err = err_out ? err_out : (int *)(b->arena + b->off_err);
    if (b->dtype == PICT_DTYPE_F32)
        hipLaunchKernelGGL(k_batch_err<float>, dim3(1), dim3(256), 0, stream, (const SolveArgs<float> *)args_dev,
                           b->count, err);
    else
        hipLaunchKernelGGL(k_batch_err<double>, dim3(1), dim3(256), 0, stream, (const SolveArgs<double> *)args_dev,
                           b->count, err);
    PICT_HIP_TRY(hipGetLastError());
    if (states_out)
        PICT_HIP_TRY(hipMemcpyAsync(states_out, b->arena + b->off_states, sizeof(RhsState) * (size_t)b->count * b->nrhs,
                                    hipMemcpyDeviceToDevice, stream));
    return PICT_OK;
}

void batch_solve_describe(const BatchSolve *b, BatchSolveInfo *out) {
    out->grid = (int)b->grid;
    out->rows_per_thread = b->rpt;
    out->rows_per_wg = b->rows_per_wg;
    out->local = b->local || b->bicg_local ? 1 : 0;
    out->jobs = (int)b->jobs.size();
    out->makespan = b->makespan;
    const bool f32 = b->dtype == PICT_DTYPE_F32;
    if (b->local)
        snprintf(out->kernel, sizeof out->kernel, "pict::k_cg_local_jobs<%s, %d, %d, false>", f32 ? "float" : "double",
                 b->rpt, b->nz);
    else if (b->bicg_local)
        snprintf(out->kernel, sizeof out->kernel, "pict::k_bicg_local_jobs<%d, %d, %d>", b->rpt, b->nz, b->nrhs);
    else
        snprintf(out->kernel, sizeof out->kernel, "pict::k_bicgstab_jobs<%s, %d>", f32 ? "float" : "double", b->nrhs);
}

}  // namespace pict

namespace {

// Synchronous solve: queue it, wait, report the states as pict_solve_info.
pict_status solve_sync(const pict_csr_desc *A, const void *rhs, void *x, int nrhs, const pict_solve_opts *o,
                       pict_solve_info *infos, hipStream_t stream) {
    if (o->use_bicg && o->bicg_with_preconditioner) {
        if (g_timing && !g_ev[0]) {
            PICT_HIP_TRY(hipEventCreate(&g_ev[0]));
            PICT_HIP_TRY(hipEventCreate(&g_ev[1]));
        }
        if (g_timing) PICT_HIP_TRY(hipEventRecord(g_ev[0], stream));
        pict_status st = pict::ilu_bicgstab_solve(A, rhs, x, nrhs, o, infos, stream);
        if (st) return st;
        if (g_timing) {
            // the events bracket the analysis copies too: the time of the whole preconditioned solve
            PICT_HIP_TRY(hipEventRecord(g_ev[1], stream));
            PICT_HIP_TRY(hipEventSynchronize(g_ev[1]));
            float ms = 0.f;
            PICT_HIP_TRY(hipEventElapsedTime(&ms, g_ev[0], g_ev[1]));
            g_last_ms = ms;
        }
        return PICT_OK;
    }
    struct Out {
        RhsState st[kMaxRhs];
        int err;
    };
    Out *dev = nullptr;
    PICT_HIP_TRY(ws_alloc((void **)&dev, sizeof(Out), stream));
    struct Guard {
        Out *p; hipStream_t s;
        ~Guard() { if (p) (void)ws_free(p, s); }
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

pict_status pict_solver_kernel_name(const pict_csr_desc *A, int32_t nrhs, const pict_solve_opts *opts, char *buf,
                                    int32_t len) {
    PICT_CHECK_ARG(A && opts && buf && len > 0, "NULL argument");
    pict_status st = validate_solve(A, nrhs, opts);
    if (st) return st;
    Plan pl;
    const bool f32 = A->dtype == PICT_DTYPE_F32;
    const char *t = f32 ? "float" : "double";
    if (opts->use_bicg && opts->bicg_with_preconditioner) {
        snprintf(buf, (size_t)len, "pict::k_bicgstab_ilu<%s>", t);
        return PICT_OK;
    }
    st = f32 ? plan_solve<float>(A, nrhs, opts, pl) : plan_solve<double>(A, nrhs, opts, pl);
    if (st) return st;
    const int nr = nrhs < kMaxRhs ? nrhs : kMaxRhs;
    if (pl.local)
        snprintf(buf, (size_t)len, "pict::k_cg_local<%s, %d, %d, %s>", t, pl.rpt, pl.nz, pl.full ? "true" : "false");
    else if (pl.bicg_local)
        snprintf(buf, (size_t)len, "pict::k_bicg_local<%d, %d, %d>", pl.rpt, pl.nz, nr);
    else if (pl.kind == Kind::kBiCG)
        snprintf(buf, (size_t)len, "pict::k_bicgstab<%s, %d>", t, nr);
    else if (pl.kind == Kind::kCGTextbook)
        snprintf(buf, (size_t)len, "pict::k_cg<%s, %d>", t, nr);
    else
        snprintf(buf, (size_t)len, "pict::k_cg1<%s, %d, %s>", t, nr, pl.lean ? "true" : "false");
    return PICT_OK;
}

pict_status pict_solve_refined(const pict_csr_desc *A, const void *rhs, double *x, const pict_solve_opts *inner,
                               double refine_tol, int32_t max_sweeps, int32_t null_space, pict_solve_info *info_out,
                               void *stream) {
    PICT_CHECK_ARG(A && rhs && x && inner && info_out, "NULL argument");
    pict_status st = validate_solve(A, 1, inner);
    if (st) return st;
    PICT_CHECK_ARG(!inner->transpose_a, "transposed refined solves are not supported");
    PICT_CHECK_ARG(refine_tol > 0.0 && max_sweeps >= 0, "refine_tol must be > 0 and max_sweeps >= 0");
    PICT_CHECK_ARG(null_space >= -1 && null_space <= 1, "null_space must be -1 (detect), 0 or 1");
    hipStream_t s = (hipStream_t)stream;
    const int n = A->n;
    const bool f32 = A->dtype == PICT_DTYPE_F32;
    const size_t es = f32 ? sizeof(float) : sizeof(double);
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t o_d = al(sizeof(double) * n), o_rr = o_d + al(es * n), o_part = o_rr + al(es * n),
                 bytes = o_part + sizeof(double) * 4 * kRefineBlocks;
    char *ws = nullptr;
    PICT_HIP_TRY(ws_alloc((void **)&ws, bytes, s));
    struct Guard {
        char *p; hipStream_t s;
        ~Guard() { if (p) (void)ws_free(p, s); }
    } guard{ws, s};
    double *r = (double *)ws, *part = (double *)(ws + o_part);
    void *d = ws + o_d, *rr_vec = ws + o_rr;
    const unsigned blocks = (unsigned)std::min<long long>(((long long)n + 255) / 256, kRefineBlocks);
    std::vector<double> hp(4 * kRefineBlocks);
    // the constant null space: given, or detected from the row sums
    bool project = null_space == 1;
    if (null_space == -1) {
        if (f32)
            hipLaunchKernelGGL(k_row_sum_ratio<float>, dim3(blocks), dim3(256), 0, s, A->row, (const float *)A->value,
                               n, part);
        else
            hipLaunchKernelGGL(k_row_sum_ratio<double>, dim3(blocks), dim3(256), 0, s, A->row,
                               (const double *)A->value, n, part);
        PICT_HIP_TRY(hipGetLastError());
        PICT_HIP_TRY(hipMemcpyAsync(hp.data(), part, sizeof(double) * blocks, hipMemcpyDeviceToHost, s));
        PICT_HIP_TRY(hipStreamSynchronize(s));
        double m = 0.0;
        for (unsigned k = 0; k < blocks; ++k) m = std::max(m, hp[k]);
        project = m < (f32 ? 1e-5 : 1e-12);
    }
    double rel = 0.0, prev = 0.0, inner_ms = 0.0, last_rms = 0.0;
    int inner_its = 0, sweep = 0;
    bool converged = false, finite = true;
    // A sweep whose single-reduction CG correction did not halve the residual (its recurrences can stall on a singular,
    // not quite symmetric FP32 matrix where textbook CG still converges) hands the remaining sweeps to the textbook
    // CG; a correction that increased the residual is taken back first.
    bool textbook = false, redo = false;
    struct Restore {
        ~Restore() { g_force_textbook = false; }
    } restore;
    pict_solve_info ii;
    for (;; ++sweep) {
        if (f32)
            hipLaunchKernelGGL(k_residual64<float>, dim3(blocks), dim3(256), 0, s, A->row, A->index,
                               (const float *)A->value, (const float *)rhs, (const double *)x, n, r, part);
        else
            hipLaunchKernelGGL(k_residual64<double>, dim3(blocks), dim3(256), 0, s, A->row, A->index,
                               (const double *)A->value, (const double *)rhs, (const double *)x, n, r, part);
        PICT_HIP_TRY(hipGetLastError());
        PICT_HIP_TRY(hipMemcpyAsync(hp.data(), part, sizeof(double) * 4 * blocks, hipMemcpyDeviceToHost, s));
        PICT_HIP_TRY(hipStreamSynchronize(s));
        double sr = 0.0, rr = 0.0, sb = 0.0, bb = 0.0;
        for (unsigned k = 0; k < blocks; ++k) {
            sr += hp[4 * k];
            rr += hp[4 * k + 1];
            sb += hp[4 * k + 2];
            bb += hp[4 * k + 3];
        }
        // with a constant null space only the part of r (and b) orthogonal to it can be solved for
        const double shift = project ? sr / n : 0.0;
        if (project) {
            rr = std::max(rr - sr * sr / n, 0.0);
            bb = std::max(bb - sb * sb / n, 0.0);
        }
        rel = bb > 0.0 ? sqrt(rr / bb) : sqrt(rr);
        finite = std::isfinite(rel);
        if (sweep > 0 && !textbook && !inner->use_bicg && !(rel <= 0.5 * prev)) {
            if (!(rel <= prev)) {  // take the correction back
                if (f32)
                    hipLaunchKernelGGL(k_refine_update<float>, dim3(blocks), dim3(256), 0, s, x, (const float *)d,
                                       -last_rms, n);
                else
                    hipLaunchKernelGGL(k_refine_update<double>, dim3(blocks), dim3(256), 0, s, x, (const double *)d,
                                       -last_rms, n);
                PICT_HIP_TRY(hipGetLastError());
            }
            textbook = g_force_textbook = true;
            redo = true;
            continue;  // the residual of x again, then the next sweep with the textbook CG
        }
        if (!finite || rel < refine_tol) {
            converged = finite;
            break;
        }
        // stop when the sweeps no longer reduce the residual (the inner solver's precision floor)
        if (sweep >= max_sweeps || (sweep > 0 && !redo && rel > 0.5 * prev)) break;
        redo = false;
        prev = rel;
        const double rms = sqrt(rr / (double)n);
        if (f32)
            hipLaunchKernelGGL(k_refine_rhs<float>, dim3(blocks), dim3(256), 0, s, (const double *)r, shift,
                               1.0 / rms, (float *)rr_vec, (float *)d, n);
        else
            hipLaunchKernelGGL(k_refine_rhs<double>, dim3(blocks), dim3(256), 0, s, (const double *)r, shift,
                               1.0 / rms, (double *)rr_vec, (double *)d, n);
        PICT_HIP_TRY(hipGetLastError());
        if ((st = solve_sync(A, rr_vec, d, 1, inner, &ii, s))) return st;
        inner_its += ii.used_iterations + 1;
        inner_ms += g_last_ms;
        last_rms = rms;
        if (f32)
            hipLaunchKernelGGL(k_refine_update<float>, dim3(blocks), dim3(256), 0, s, x, (const float *)d, rms, n);
        else
            hipLaunchKernelGGL(k_refine_update<double>, dim3(blocks), dim3(256), 0, s, x, (const double *)d, rms, n);
        PICT_HIP_TRY(hipGetLastError());
    }
    if (g_timing) g_last_ms = inner_ms;  // the kernel time of the refined solve: the sum over its correction solves
    info_out->final_residual = rel;
    info_out->used_iterations = inner_its;
    info_out->converged = converged ? 1 : 0;
    info_out->is_finite_residual = finite ? 1 : 0;
    info_out->reserved = sweep;
    return PICT_OK;
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
    PICT_HIP_TRY(ws_alloc((void **)&missing, sizeof(int), s));
    struct Guard {
        int *p; hipStream_t s;
        ~Guard() { if (p) (void)ws_free(p, s); }
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
