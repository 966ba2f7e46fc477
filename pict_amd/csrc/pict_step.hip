// pict_step.hip -- a whole PISO time step queued natively on one stream.
//
// The orthogonal branch of the reference driver (PISOtorch_simulation.py:1054-1327:
// advection matrix and RHS, BiCGStab/CG velocity solve, `corrector_steps` x {pressure
// matrix, RHS, divergence, pressure solve, mean subtraction, copy to blocks, velocity
// correction}, copy to blocks) with every stage and solve queued back to back: the
// host waits once per step, to read the solvers' final states, instead of once per
// solve, and no interpreter runs between the stages.
#include <mutex>
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <vector>

#include "pict_host.hpp"
#include "pict_solver.hpp"

namespace pict {

template <typename R>
__global__ __launch_bounds__(256) void k_fill(R *v, size_t n, R value) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        v[i] = value;
}

// FP64 partial sums of v, one per workgroup (fixed order: deterministic)
template <typename R>
__global__ __launch_bounds__(kMeanThreads) void k_partial_sum(const R *v, int n, double *part) {
    __shared__ double lds[kMeanThreads / 64];
    double s = 0.0;
    for (int i = blockIdx.x * kMeanThreads + threadIdx.x; i < n; i += gridDim.x * kMeanThreads) s += (double)v[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kMeanThreads / 64; ++w) t += lds[w];
        part[blockIdx.x] = t;
    }
}

// v -= mean(v), the mean from the partial sums (every workgroup sums them in the same order)
template <typename R>
__global__ __launch_bounds__(kMeanThreads) void k_sub_mean(R *v, int n, const double *part, int nparts) {
    __shared__ double mean;
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int b = 0; b < nparts; ++b) t += part[b];
        mean = t / (double)n;
    }
    __syncthreads();
    const R m = (R)mean;
    for (int i = blockIdx.x * kMeanThreads + threadIdx.x; i < n; i += gridDim.x * kMeanThreads) v[i] -= m;
}

template <typename R>
pict_status fill(void *v, size_t n, double value, hipStream_t s) {
    const unsigned blocks = (unsigned)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024);
    hipLaunchKernelGGL(k_fill<R>, dim3(blocks ? blocks : 1), dim3(256), 0, s, (R *)v, n, (R)value);
    return check_launch();
}

template <typename R>
pict_status subtract_mean(void *v, int n, double *part, hipStream_t s) {
    hipLaunchKernelGGL(k_partial_sum<R>, dim3(kMeanBlocks), dim3(kMeanThreads), 0, s, (const R *)v, n, part);
    hipLaunchKernelGGL(k_sub_mean<R>, dim3(kMeanBlocks), dim3(kMeanThreads), 0, s, (R *)v, n, (const double *)part,
                       kMeanBlocks);
    return check_launch();
}

// y += alpha * x
template <typename R>
__global__ __launch_bounds__(256) void k_axpy(R *y, const R *x, size_t n, R alpha) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] += alpha * x[i];
}

// out[e] += alpha * a[i] * b[j] for every pattern entry e = (i, j) (the SDDMM of the solve VJP)
template <typename R>
__global__ __launch_bounds__(256) void k_outer_acc(R *out, const R *a, const R *b, const int32_t *row,
                                                    const int32_t *col, int n, R alpha) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const R ai = alpha * a[i];
    for (int e = row[i]; e < row[i + 1]; ++e) out[e] += ai * b[col[e]];
}

// Block gradient fields <-> global vectors: mode 0: gv += block field; mode 1: block field = 0.
// field 2: velocity_grad (D components), 3: pressure_grad.
template <typename R>
__global__ __launch_bounds__(kCellsPerTile) void k_block_grads(DomainK<R> dom, R *gv, int field, int mode) {
    const pict_block_desc *b;
    int bi, f;
    if (!tile_cell(dom, b, bi, f)) return;
    R *bv = reinterpret_cast<R *>(field == 2 ? b->velocity_grad : b->pressure_grad);
    if (!bv) return;
    const int g = b->global_offset + f;
    const int comps = field == 2 ? dom.D : 1;
    for (int d = 0; d < comps; ++d) {
        if (mode == 0) gv[(size_t)d * dom.N + g] += bv[(size_t)d * b->stride[3] + f];
        else bv[(size_t)d * b->stride[3] + f] = R(0);
    }
}

template <typename R>
pict_status axpy(void *y, const void *x, size_t n, double alpha, hipStream_t s) {
    const unsigned blocks = (unsigned)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024);
    hipLaunchKernelGGL(k_axpy<R>, dim3(blocks ? blocks : 1), dim3(256), 0, s, (R *)y, (const R *)x, n, (R)alpha);
    return check_launch();
}

// solver records of the last native step of this thread (kernel times for diagnostics)
struct StepTiming {
    std::vector<hipEvent_t> ev;
    std::vector<double> ms;
    bool enabled = false;
};
thread_local StepTiming g_step_timing;

pict_status timing_events(int count, hipEvent_t *&out) {
    StepTiming &t = g_step_timing;
    out = nullptr;
    if (!t.enabled) return PICT_OK;
    while ((int)t.ev.size() < 2 * count) {
        hipEvent_t e;
        PICT_HIP_TRY(hipEventCreate(&e));
        t.ev.push_back(e);
    }
    out = t.ev.data();
    return PICT_OK;
}

}  // namespace pict

using namespace pict;

namespace {

pict_csr_desc csr_of(const pict_atlas *a, void *values) {
    pict_csr_desc m;
    m.n = a->desc.num_cells;
    m.nnz = a->desc.nnz;
    m.dtype = a->desc.dtype;
    m.reserved = 0;
    m.row = a->desc.csr_row;
    m.index = a->desc.csr_index;
    m.value = values;
    m.transpose_perm = nullptr;
    return m;
}

// pinned host staging for the solver states of one step (grows as needed, kept per thread)
struct HostStates {
    void *p = nullptr;
    size_t bytes = 0;
    ~HostStates() {
        if (p) (void)hipHostFree(p);
    }
};
thread_local HostStates g_host_states;

}  // namespace

static pict_status copy_dev(void *dst, const void *src, size_t bytes, hipStream_t s) {
    PICT_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
    return PICT_OK;
}

namespace {

// Device and host places of one native step or reverse pass.
//   dev:  per solve a slot [kMaxRhs states | error flag]; the mean partial sums; for the
//         reverse pass its vectors g_u, A_tot, C_tot, lambda.
//   solve_ws[i]: workspace of solve i (null: from the library pool).
//   host: pinned destination of the solve slots (copied at the end of the queued work).
struct Work {
    char *dev = nullptr;
    void *solve_ws[65] = {};
    size_t solve_ws_bytes[65] = {};
    void *host = nullptr;
    hipEvent_t *ev = nullptr;
};

struct Sizes {
    size_t slot, o_part, o_gu, o_at, o_ct, o_lam, dev_bytes;
    int nsolves;
};

Sizes sizes_of(const pict_atlas *a, const pict_step_opts *o, bool vjp) {
    const pict_domain_desc &d = a->desc;
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t es = d.dtype == PICT_DTYPE_F32 ? sizeof(float) : sizeof(double);
    const size_t vecD = es * d.num_dims * d.num_cells, vec1 = es * d.num_cells, nzb = es * d.nnz;
    Sizes z;
    z.nsolves = 1 + o->corrector_steps;
    z.slot = al(4 * solve_state_bytes() + sizeof(int));
    z.o_part = z.slot * z.nsolves;
    z.o_gu = z.o_part + al(sizeof(double) * kMeanBlocks);
    if (!vjp) {
        z.o_at = z.o_ct = z.o_lam = 0;
        z.dev_bytes = z.o_gu;
        return z;
    }
    z.o_at = z.o_gu + al(vecD);
    z.o_ct = z.o_at + al(vec1);
    z.o_lam = z.o_ct + al(nzb);
    z.dev_bytes = z.o_lam + al(vecD);
    return z;
}

// the solves of a step (vjp: transposed, zero-RHS shortcut): i = 0 advection, 1.. pressure
void solve_systems(const pict_atlas *a, const pict_step_opts *o, bool vjp, pict_csr_desc &C, pict_csr_desc &P,
                   pict_solve_opts &adv, pict_solve_opts &prs) {
    C = csr_of(a, a->desc.C_value);
    P = csr_of(a, a->desc.P_value);
    adv = o->advection;
    prs = o->pressure;
    if (vjp) {
        C.transpose_perm = a->tperm;
        P.transpose_perm = a->tperm;
        adv.transpose_a = prs.transpose_a = 1;
        adv.zero_rhs_shortcut = prs.zero_rhs_shortcut = 1;
    }
}

// the forward step queued on s (no allocation, no host wait when w's workspaces are given)
pict_status enqueue_step(const pict_atlas *a, const pict_step_opts *o, const pict_step_tape *tape, const Work &w,
                         const Sizes &z, hipStream_t s) {
    const pict_domain_desc &d = a->desc;
    void *stream = s;
    const int D = d.num_dims, N = d.num_cells, K = o->corrector_steps;
    const bool f32 = d.dtype == PICT_DTYPE_F32;
    const size_t sb = solve_state_bytes();
    char *dev = w.dev;
    double *part = (double *)(dev + z.o_part);
    hipEvent_t *ev = w.ev;
    pict_status st;
    const double dt = o->time_step;
    const size_t es = f32 ? sizeof(float) : sizeof(double);
    const size_t vecD = es * D * N, vec1 = es * N, nzb = es * d.nnz;
    pict_csr_desc C, P;
    pict_solve_opts adv, prs;
    solve_systems(a, o, false, C, P, adv, prs);
    // ---- advection (PISOtorch_simulation.py:1161-1188): matrix, RHS and start vector in one launch ----
    if ((st = stage_advection_setup(a, dt, s))) return st;
    if (tape && ((st = copy_dev(tape->C_value, d.C_value, nzb, s)) || (st = copy_dev(tape->A, d.A, vec1, s))))
        return st;
    if ((st = solve_enqueue(&C, d.velocity_rhs, d.velocity_result, D, &adv, dev, (int *)(dev + 4 * sb),
                            ev ? ev[0] : nullptr, ev ? ev[1] : nullptr, s, w.solve_ws[0], w.solve_ws_bytes[0])))
        return st;
    if (tape && (st = copy_dev(tape->velocity_star, d.velocity_result, vecD, s))) return st;
    // ---- correctors (:1229-1320) ----
    for (int k = 0; k < K; ++k) {
        char *sl = dev + z.slot * (1 + k);
        // pressure matrix (first corrector: it depends on A only), RHS and start vector in one launch; divergence
        if ((st = stage_pressure_setup(a, dt, k == 0, s))) return st;
        if ((st = pict_setup_pressure_rhs_div(a, dt, 0, 0, o->time_step_norm, stream))) return st;
        if (tape && k == 0 && (st = copy_dev(tape->P_value, d.P_value, nzb, s))) return st;
        if (tape && (st = copy_dev((char *)tape->pressure_rhs + vecD * k, d.pressure_rhs, vecD, s))) return st;
        if ((st = solve_enqueue(&P, d.pressure_rhs_div, d.pressure_result, 1, &prs, sl, (int *)(sl + 4 * sb),
                                ev ? ev[2 * (1 + k)] : nullptr, ev ? ev[2 * (1 + k) + 1] : nullptr, s,
                                w.solve_ws[1 + k], w.solve_ws_bytes[1 + k])))
            return st;
        if (tape && (st = copy_dev((char *)tape->pressure_raw + vec1 * k, d.pressure_result, vec1, s))) return st;
        if (o->normalize_pressure) {
            if ((st = f32 ? subtract_mean<float>(d.pressure_result, N, part, s)
                          : subtract_mean<double>(d.pressure_result, N, part, s)))
                return st;
        }
        if ((st = pict_copy_pressure_result_to_blocks(a, stream))) return st;
        if ((st = pict_correct_velocity(a, dt, o->velocity_corrector, o->time_step_norm, stream))) return st;
        if (tape && (st = copy_dev((char *)tape->velocity_result + vecD * k, d.velocity_result, vecD, s))) return st;
    }
    if ((st = pict_copy_velocity_result_to_blocks(a, stream))) return st;
    PICT_HIP_TRY(hipMemcpyAsync(w.host, dev, z.slot * z.nsolves, hipMemcpyDeviceToHost, s));
    return PICT_OK;
}

// the reverse pass queued on s (the order in which PISOtorch_diff's autograd graph visits the
// stages of the orthogonal step, PISOtorch_diff.py:295-1234)
pict_status enqueue_vjp(const pict_atlas *a, const pict_step_opts *o, const pict_step_tape *tape, void *velocity_grad,
                        const Work &w, const Sizes &z, hipStream_t s) {
    const pict_domain_desc &d = a->desc;
    void *stream = s;
    const int D = d.num_dims, N = d.num_cells, K = o->corrector_steps;
    const bool f32 = d.dtype == PICT_DTYPE_F32;
    const size_t es = f32 ? sizeof(float) : sizeof(double);
    const size_t vecD = es * D * N, vec1 = es * N, nzb = es * d.nnz;
    const size_t sb = solve_state_bytes();
    char *dev = w.dev;
    double *part = (double *)(dev + z.o_part);
    void *g_u = dev + z.o_gu, *A_tot = dev + z.o_at, *C_tot = dev + z.o_ct, *lam = dev + z.o_lam;
    const double dt = o->time_step;
    hipEvent_t *ev = w.ev;
    pict_status st;
#define PICT_TRY(expr) do { if ((st = (expr))) return st; } while (0)
#define PICT_T(fn, ...) (f32 ? fn<float>(__VA_ARGS__) : fn<double>(__VA_ARGS__))
    auto block_grads = [&](void *gv, int field, int mode) -> pict_status {
        if (f32) PICT_LAUNCH_CELLS(a, stream, (k_block_grads<float>), a->view<float>(), (float *)gv, field, mode);
        else PICT_LAUNCH_CELLS(a, stream, (k_block_grads<double>), a->view<double>(), (double *)gv, field, mode);
        return check_launch();
    };
    auto outer_acc = [&](void *out, const void *av, const void *bv, double alpha) -> pict_status {
        const unsigned blocks = (unsigned)((N + 255) / 256);
        if (f32)
            hipLaunchKernelGGL(k_outer_acc<float>, dim3(blocks), dim3(256), 0, s, (float *)out, (const float *)av,
                               (const float *)bv, d.csr_row, d.csr_index, N, (float)alpha);
        else
            hipLaunchKernelGGL(k_outer_acc<double>, dim3(blocks), dim3(256), 0, s, (double *)out, (const double *)av,
                               (const double *)bv, d.csr_row, d.csr_index, N, alpha);
        return check_launch();
    };
    pict_csr_desc C, P;
    pict_solve_opts adv, prs;
    solve_systems(a, o, true, C, P, adv, prs);

    // forward state the VJP kernels read
    PICT_TRY(copy_dev(d.C_value, tape->C_value, nzb, s));
    PICT_TRY(copy_dev(d.A, tape->A, vec1, s));
    PICT_TRY(copy_dev(d.P_value, tape->P_value, nzb, s));
    PICT_TRY(PICT_T(fill, g_u, (size_t)D * N, 0.0, s));
    PICT_TRY(PICT_T(fill, A_tot, (size_t)N, 0.0, s));
    PICT_TRY(PICT_T(fill, C_tot, (size_t)d.nnz, 0.0, s));
    // CopyVelocityResultToBlocks: the output block-velocity gradient is the velocity-result gradient
    PICT_TRY(copy_dev(d.velocity_result_grad, velocity_grad, vecD, s));
    for (int k = K - 1; k >= 0; --k) {
        char *sl = dev + z.slot * (1 + k);
        // ---- CorrectVelocity (inputs A, block pressure p^_k, pressure RHS) ----
        PICT_TRY(copy_dev(d.pressure_result, (const char *)tape->pressure_raw + vec1 * k, vec1, s));
        if (o->normalize_pressure) PICT_TRY(PICT_T(subtract_mean, d.pressure_result, N, part, s));
        PICT_TRY(pict_copy_pressure_result_to_blocks(a, stream));
        PICT_TRY(block_grads(nullptr, 3, 1));
        PICT_TRY(pict_correct_velocity_grad(a, dt, o->time_step_norm, stream));
        PICT_TRY(PICT_T(axpy, A_tot, d.A_grad, (size_t)N, 1.0, s));
        // ---- CopyPressureResultToBlocks, mean subtraction, pressure solve ----
        PICT_TRY(pict_copy_pressure_result_grad_from_blocks(a, stream));
        if (o->normalize_pressure) PICT_TRY(PICT_T(subtract_mean, d.pressure_result_grad, N, part, s));
        PICT_TRY(PICT_T(fill, d.pressure_rhs_div_grad, (size_t)N, 0.0, s));
        PICT_TRY(solve_enqueue(&P, d.pressure_result_grad, d.pressure_rhs_div_grad, 1, &prs, sl,
                               (int *)(sl + 4 * sb), ev ? ev[2 * (1 + k)] : nullptr,
                               ev ? ev[2 * (1 + k) + 1] : nullptr, s, w.solve_ws[1 + k], w.solve_ws_bytes[1 + k]));
        PICT_TRY(PICT_T(fill, d.P_grad, (size_t)d.nnz, 0.0, s));
        PICT_TRY(outer_acc(d.P_grad, d.pressure_rhs_div_grad, (const char *)tape->pressure_raw + vec1 * k, -1.0));
        // ---- SetupPressureCorrection (inputs velocity, source, velocity result, boundaries, viscosity, A, C) ----
        PICT_TRY(copy_dev(d.velocity_result, k == 0 ? tape->velocity_star
                                                     : (const char *)tape->velocity_result + vecD * (k - 1), vecD, s));
        PICT_TRY(copy_dev(d.pressure_rhs, (const char *)tape->pressure_rhs + vecD * k, vecD, s));
        PICT_TRY(PICT_T(fill, d.velocity_result_grad, (size_t)D * N, 0.0, s));
        PICT_TRY(PICT_T(fill, d.A_grad, (size_t)N, 0.0, s));
        PICT_TRY(pict_setup_pressure_correction_grad(a, dt, 0, 0, o->time_step_norm, stream));
        PICT_TRY(PICT_T(axpy, A_tot, d.A_grad, (size_t)N, 1.0, s));
        PICT_TRY(PICT_T(axpy, C_tot, d.C_grad, (size_t)d.nnz, 1.0, s));
        PICT_TRY(block_grads(g_u, 2, 0));
    }
    // ---- advection solve C u* = rhs: lambda = C^-T g, C_grad -= lambda u*^T ----
    PICT_TRY(PICT_T(fill, lam, (size_t)D * N, 0.0, s));
    PICT_TRY(solve_enqueue(&C, d.velocity_result_grad, lam, D, &adv, dev, (int *)(dev + 4 * sb),
                           ev ? ev[0] : nullptr, ev ? ev[1] : nullptr, s, w.solve_ws[0], w.solve_ws_bytes[0]));
    for (int c = 0; c < D; ++c)
        PICT_TRY(outer_acc(C_tot, (const char *)lam + vec1 * c, (const char *)tape->velocity_star + vec1 * c, -1.0));
    // ---- SetupAdvectionVelocity (rhs gradient = lambda) ----
    PICT_TRY(copy_dev(d.velocity_rhs_grad, lam, vecD, s));
    PICT_TRY(pict_setup_advection_velocity_grad(a, dt, 0, 0, stream));
    PICT_TRY(block_grads(g_u, 2, 0));
    // ---- SetupAdvectionMatrix (A and C gradients of every consumer) ----
    PICT_TRY(copy_dev(d.A_grad, A_tot, vec1, s));
    PICT_TRY(copy_dev(d.C_grad, C_tot, nzb, s));
    PICT_TRY(block_grads(nullptr, 2, 1));
    PICT_TRY(pict_setup_advection_matrix_grad(a, dt, 0, stream));
    PICT_TRY(block_grads(g_u, 2, 0));
    PICT_TRY(copy_dev(velocity_grad, g_u, vecD, s));
#undef PICT_TRY
#undef PICT_T
    PICT_HIP_TRY(hipMemcpyAsync(w.host, dev, z.slot * z.nsolves, hipMemcpyDeviceToHost, s));
    return PICT_OK;
}

// solver infos from the host copy of the solve slots
pict_status collect(const void *host, const Sizes &z, const pict_step_opts *o, int D, pict_solve_info *infos_out,
                    hipEvent_t *ev) {
    const char *hp = (const char *)host;
    const size_t sb = solve_state_bytes();
    int err = 0;
    for (int i = 0; i < z.nsolves; ++i) err |= *(const int *)(hp + z.slot * i + 4 * sb);
    if (err) {
        set_error("solver grid barrier exceeded its spin limit");
        return PICT_ERR_DEVICE_SYNC;
    }
    solve_state_to_info(hp, D, infos_out);
    for (int k = 0; k < o->corrector_steps; ++k) solve_state_to_info(hp + z.slot * (1 + k), 1, infos_out + D + k);
    if (ev) {
        g_step_timing.ms.assign(z.nsolves, 0.0);
        for (int i = 0; i < z.nsolves; ++i) {
            float ms = 0.f;
            PICT_HIP_TRY(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
            g_step_timing.ms[i] = ms;
        }
    }
    return PICT_OK;
}

// ---------------------------------------------------------------------------
// Graphs: the queued work of a step or reverse pass is captured once and replayed.  Its
// short kernels (2-30 us) otherwise wait on the host's launches (about 10 us each); a
// replay dispatches them about 2 us apart.  Everything a capture bakes in is in the key:
// the domain-level descriptor (vector addresses, sizes, viscosity), the options, the tape
// and gradient addresses and the transpose map.  Block tensors are read through the
// atlas's device block table at run time, so swapping them needs no new capture.  Each
// graph owns its workspaces (one device arena) and the pinned buffer of its solver states;
// every replay is followed by a host wait, so an entry can be replaced at any call.
// ---------------------------------------------------------------------------
}  // namespace

namespace pict {

struct StepKey {
    int kind;  // 0 step, 1 step recording a tape, 2 reverse pass
    int device;
    pict_domain_desc d;
    pict_step_opts o;
    pict_step_tape t;
    const void *vgrad;
    const int32_t *tperm;
};

struct GraphEntry {
    bool valid = false;
    StepKey key;
    hipGraphExec_t exec = nullptr;
    void *arena = nullptr;
    void *host = nullptr;
    Sizes z;
    void release() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (arena) (void)hipFree(arena);
        if (host) (void)hipHostFree(host);
        exec = nullptr;
        arena = host = nullptr;
        valid = false;
    }
};

struct GraphCache {
    GraphEntry e[3];
    StepKey miss[3];       // the last key of each kind that had no graph
    bool miss_valid[3] = {};
    hipStream_t capture = nullptr;
    // device workspace of the directly queued steps of each kind (step slots, reverse-pass vectors and one
    // workspace per solve), kept between calls: the library pool's per-solve allocations cost host time
    void *arena[3] = {};
    size_t arena_bytes[3] = {};
    ~GraphCache() {
        for (GraphEntry &g : e) g.release();
        if (capture) (void)hipStreamDestroy(capture);
        for (void *p : arena)
            if (p) (void)hipFree(p);
    }
};

}  // namespace pict

namespace {

bool graphs_enabled() {
    const char *e = getenv("PICT_STEP_GRAPHS");
    return !e || atoi(e) != 0;
}

StepKey key_of(const pict_atlas *a, int kind, const pict_step_opts *o, const pict_step_tape *tape, const void *vgrad) {
    StepKey k;
    memset(&k, 0, sizeof k);
    k.kind = kind;
    k.device = a->device;
    k.d = a->desc;
    k.d.blocks = nullptr;
    k.o = *o;
    if (tape) k.t = *tape;
    k.vgrad = vgrad;
    k.tperm = kind == 2 ? a->tperm : nullptr;
    return k;
}

// The graph of `kind` for this call; null when the key is new: a key is captured when it comes
// twice in a row, so calls whose addresses change every time (a new tape per step) run directly.
pict_status graph_for(const pict_atlas *a, int kind, const pict_step_opts *o, const pict_step_tape *tape,
                      void *vgrad, GraphEntry *&out) {
    out = nullptr;
    if (!a->step_graphs) a->step_graphs = new GraphCache();
    GraphCache &c = *a->step_graphs;
    GraphEntry &g = c.e[kind];
    const StepKey key = key_of(a, kind, o, tape, vgrad);
    if (g.valid && memcmp(&g.key, &key, sizeof key) == 0) {
        out = &g;
        return PICT_OK;
    }
    if (!(c.miss_valid[kind] && memcmp(&c.miss[kind], &key, sizeof key) == 0)) {
        c.miss[kind] = key;
        c.miss_valid[kind] = true;
        return PICT_OK;
    }
    c.miss_valid[kind] = false;
    g.release();
    if (!c.capture) PICT_HIP_TRY(hipStreamCreateWithFlags(&c.capture, hipStreamNonBlocking));
    const bool vjp = kind == 2;
    const Sizes z = sizes_of(a, o, vjp);
    pict_csr_desc C, P;
    pict_solve_opts adv, prs;
    solve_systems(a, o, vjp, C, P, adv, prs);
    size_t wsb[65];
    size_t total = z.dev_bytes;
    pict_status st = solve_workspace_bytes(&C, a->desc.num_dims, &adv, &wsb[0]);
    if (st) return st;
    for (int k = 0; k < o->corrector_steps; ++k)
        if ((st = solve_workspace_bytes(&P, 1, &prs, &wsb[1 + k]))) return st;
    for (int i = 0; i < z.nsolves; ++i) total += (wsb[i] + 255) & ~size_t(255);
    PICT_HIP_TRY(hipMalloc(&g.arena, total));
    PICT_HIP_TRY(hipHostMalloc(&g.host, z.slot * z.nsolves, hipHostMallocDefault));
    Work w;
    w.dev = (char *)g.arena;
    char *p = w.dev + z.dev_bytes;
    for (int i = 0; i < z.nsolves; ++i) {
        w.solve_ws[i] = p;
        w.solve_ws_bytes[i] = wsb[i];
        p += (wsb[i] + 255) & ~size_t(255);
    }
    w.host = g.host;
    hipGraph_t graph = nullptr;
    PICT_HIP_TRY(hipStreamBeginCapture(c.capture, hipStreamCaptureModeThreadLocal));
    st = vjp ? enqueue_vjp(a, o, tape, vgrad, w, z, c.capture) : enqueue_step(a, o, tape, w, z, c.capture);
    const hipError_t ce = hipStreamEndCapture(c.capture, &graph);
    if (st || ce != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        g.release();
        if (st) return st;
        set_error("step capture failed: %s", hipGetErrorString(ce));
        return PICT_ERR_HIP;
    }
    const hipError_t ie = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
        g.release();
        set_error("step graph instantiation failed: %s", hipGetErrorString(ie));
        return PICT_ERR_HIP;
    }
    g.key = key;
    g.z = z;
    g.valid = true;
    out = &g;
    return PICT_OK;
}

// The kept workspace of `kind` laid out for this call: w.dev (z.dev_bytes) and one workspace per solve.  Work
// queued earlier on the atlas's stream may still use it; it is replaced (after a device wait) only when a call
// needs more.
pict_status direct_arena(const pict_atlas *a, int kind, const pict_step_opts *o, const Sizes &z, Work &w) {
    if (!a->step_graphs) a->step_graphs = new GraphCache();
    GraphCache &c = *a->step_graphs;
    const bool vjp = kind == 2;
    pict_csr_desc C, P;
    pict_solve_opts adv, prs;
    solve_systems(a, o, vjp, C, P, adv, prs);
    size_t wsb[65];
    pict_status st = solve_workspace_bytes(&C, a->desc.num_dims, &adv, &wsb[0]);
    if (st) return st;
    for (int k = 0; k < o->corrector_steps; ++k)
        if ((st = solve_workspace_bytes(&P, 1, &prs, &wsb[1 + k]))) return st;
    size_t total = (z.dev_bytes + 255) & ~size_t(255);
    for (int i = 0; i < z.nsolves; ++i) total += (wsb[i] + 255) & ~size_t(255);
    if (c.arena_bytes[kind] < total) {
        if (c.arena[kind]) {
            PICT_HIP_TRY(hipDeviceSynchronize());
            PICT_HIP_TRY(hipFree(c.arena[kind]));
            c.arena[kind] = nullptr;
            c.arena_bytes[kind] = 0;
        }
        PICT_HIP_TRY(hipMalloc(&c.arena[kind], total));
        c.arena_bytes[kind] = total;
    }
    w.dev = (char *)c.arena[kind];
    char *p = w.dev + ((z.dev_bytes + 255) & ~size_t(255));
    for (int i = 0; i < z.nsolves; ++i) {
        w.solve_ws[i] = p;
        w.solve_ws_bytes[i] = wsb[i];
        p += (wsb[i] + 255) & ~size_t(255);
    }
    return PICT_OK;
}

// one step or reverse pass: replayed graph, or (timing, graphs off) queued directly
pict_status run(const pict_atlas *a, int kind, const pict_step_opts *o, const pict_step_tape *tape, void *vgrad,
                pict_solve_info *infos_out, hipStream_t s) {
    const bool vjp = kind == 2;
    hipEvent_t *ev = nullptr;
    const Sizes z = sizes_of(a, o, vjp);
    pict_status st = timing_events(z.nsolves, ev);
    if (st) return st;
    if (!ev && graphs_enabled()) {
        GraphEntry *g = nullptr;
        if ((st = graph_for(a, kind, o, tape, vgrad, g))) return st;
        if (g) {
            PICT_HIP_TRY(hipGraphLaunch(g->exec, s));
            PICT_HIP_TRY(hipStreamSynchronize(s));
            return collect(g->host, g->z, o, a->desc.num_dims, infos_out, nullptr);
        }
    }
    Work w;
    if ((st = direct_arena(a, kind, o, z, w))) return st;
    HostStates &h = g_host_states;
    if (h.bytes < z.slot * z.nsolves) {
        if (h.p) PICT_HIP_TRY(hipHostFree(h.p));
        h.p = nullptr;
        PICT_HIP_TRY(hipHostMalloc(&h.p, z.slot * z.nsolves, hipHostMallocDefault));
        h.bytes = z.slot * z.nsolves;
    }
    w.host = h.p;
    w.ev = ev;
    st = vjp ? enqueue_vjp(a, o, tape, vgrad, w, z, s) : enqueue_step(a, o, tape, w, z, s);
    if (st) return st;
    PICT_HIP_TRY(hipStreamSynchronize(s));
    return collect(h.p, z, o, a->desc.num_dims, infos_out, ev);
}

// Queued steps whose solver states are still on their way to the host: a ring of pinned slots, each with the
// event recorded after the step's last copy.  The caller collects a slot with pict_step_wait.  One ring per
// process (a reverse pass may be queued from the autograd engine's thread and collected from another), kept
// until the process ends.
constexpr int kPendingSteps = 16;
struct PendingStep {
    bool busy = false;
    hipEvent_t ev = nullptr;
    void *host = nullptr;
    size_t bytes = 0;
    Sizes z{};
    pict_step_opts o{};
    int D = 0;
};
struct PendingRing {
    std::mutex mu;
    PendingStep slot[kPendingSteps];
};
PendingRing &pending_ring() {
    static PendingRing *r = new PendingRing;
    return *r;
}

pict_status run_async(const pict_atlas *a, int kind, const pict_step_opts *o, const pict_step_tape *tape,
                      void *vgrad, int32_t *ticket, hipStream_t s) {
    const bool vjp = kind == 2;
    const Sizes z = sizes_of(a, o, vjp);
    if (g_step_timing.enabled) {
        set_error("an asynchronous step cannot be timed (solver timing is on)");
        return PICT_ERR_INVALID;
    }
    PendingRing &ring = pending_ring();
    int idx = -1;
    {
        std::lock_guard<std::mutex> lock(ring.mu);
        for (int i = 0; i < kPendingSteps; ++i)
            if (!ring.slot[i].busy) {
                idx = i;
                ring.slot[i].busy = true;  // reserved; released below if queueing fails
                break;
            }
    }
    if (idx < 0) {
        set_error("%d asynchronous steps are pending: collect them with pict_step_wait", kPendingSteps);
        return PICT_ERR_INVALID;
    }
    PendingStep &p = ring.slot[idx];
    struct Release {
        PendingRing &r; PendingStep &p; bool keep = false;
        ~Release() { if (!keep) { std::lock_guard<std::mutex> lock(r.mu); p.busy = false; } }
    } release{ring, p};
    const size_t need = z.slot * z.nsolves;
    if (p.bytes < need) {
        if (p.host) PICT_HIP_TRY(hipHostFree(p.host));
        p.host = nullptr;
        p.bytes = 0;
        PICT_HIP_TRY(hipHostMalloc(&p.host, need, hipHostMallocDefault));
        p.bytes = need;
    }
    if (!p.ev) PICT_HIP_TRY(hipEventCreateWithFlags(&p.ev, hipEventDisableTiming));
    Work w;
    pict_status sa = direct_arena(a, kind, o, z, w);
    if (sa) return sa;
    w.host = p.host;
    w.ev = nullptr;
    pict_status st = vjp ? enqueue_vjp(a, o, tape, vgrad, w, z, s) : enqueue_step(a, o, tape, w, z, s);
    if (st) return st;
    PICT_HIP_TRY(hipEventRecord(p.ev, s));
    p.z = z;
    p.o = *o;
    p.D = a->desc.num_dims;
    release.keep = true;
    *ticket = idx;
    return PICT_OK;
}

}  // namespace

namespace pict {
void step_graphs_free(pict_atlas *a) {
    delete a->step_graphs;
    a->step_graphs = nullptr;
}
}  // namespace pict

extern "C" {

static pict_status step_entry(const pict_atlas *a, const pict_step_opts *o, const pict_step_tape *tape,
                              pict_solve_info *infos_out, int32_t *ticket, void *stream) {
    PICT_CHECK_ARG(a && o && (infos_out || ticket), "NULL argument");
    const pict_domain_desc &d = a->desc;
    PICT_CHECK_ARG(d.C_value && d.P_value && d.A && d.velocity_rhs && d.velocity_result && d.pressure_rhs &&
                       d.pressure_rhs_div && d.pressure_result,
                   "domain tensors missing");
    PICT_CHECK_ARG(o->corrector_steps >= 0 && o->corrector_steps <= 64, "corrector_steps must be 0..64");
    PICT_CHECK_ARG(o->advection.transpose_a == 0 && o->pressure.transpose_a == 0, "transposed solves in a step");
    if (tape)
        PICT_CHECK_ARG(tape->C_value && tape->A && tape->P_value && tape->velocity_star &&
                           (o->corrector_steps == 0 || (tape->pressure_rhs && tape->pressure_raw && tape->velocity_result)),
                       "tape buffers missing");
    if (ticket) return run_async(a, tape ? 1 : 0, o, tape, nullptr, ticket, (hipStream_t)stream);
    return run(a, tape ? 1 : 0, o, tape, nullptr, infos_out, (hipStream_t)stream);
}

pict_status pict_piso_step(const pict_atlas *a, const pict_step_opts *o, const pict_step_tape *tape,
                           pict_solve_info *infos_out, void *stream) {
    PICT_CHECK_ARG(infos_out, "NULL argument");
    return step_entry(a, o, tape, infos_out, nullptr, stream);
}

pict_status pict_piso_step_async(const pict_atlas *a, const pict_step_opts *o, const pict_step_tape *tape,
                                 int32_t *ticket, void *stream) {
    PICT_CHECK_ARG(ticket, "NULL argument");
    return step_entry(a, o, tape, nullptr, ticket, stream);
}

// Reverse pass of one pict_piso_step (the adjoint of the orthogonal step).
static pict_status vjp_entry(const pict_atlas *a, const pict_step_opts *o, const pict_step_tape *tape,
                             void *velocity_grad, pict_solve_info *infos_out, int32_t *ticket, void *stream) {
    PICT_CHECK_ARG(a && o && tape && velocity_grad && (infos_out || ticket), "NULL argument");
    const pict_domain_desc &d = a->desc;
    PICT_CHECK_ARG(d.C_value && d.P_value && d.A && d.velocity_rhs && d.velocity_result && d.pressure_rhs &&
                       d.pressure_rhs_div && d.pressure_result,
                   "domain tensors missing");
    PICT_CHECK_ARG(d.C_grad && d.P_grad && d.A_grad && d.velocity_rhs_grad && d.velocity_result_grad &&
                       d.pressure_rhs_grad && d.pressure_rhs_div_grad && d.pressure_result_grad,
                   "domain gradient tensors missing");
    for (int i = 0; i < d.num_blocks; ++i)
        PICT_CHECK_ARG(a->host_blocks[i].velocity_grad && a->host_blocks[i].pressure_grad,
                       "block %d: velocity/pressure gradient missing", i);
    PICT_CHECK_ARG(o->corrector_steps >= 0 && o->corrector_steps <= 64, "corrector_steps must be 0..64");
    PICT_CHECK_ARG(o->velocity_corrector == 0 || o->velocity_corrector == 1, "velocity corrector must be 0 or 1");
    PICT_CHECK_ARG(tape->C_value && tape->A && tape->P_value && tape->velocity_star &&
                       (o->corrector_steps == 0 || (tape->pressure_rhs && tape->pressure_raw && tape->velocity_result)),
                   "tape buffers missing");
    hipStream_t s = (hipStream_t)stream;
    // transposed matrices share the pattern's A^T map, built on the atlas's first reverse pass
    if (!a->tperm || a->tperm_row != d.csr_row || a->tperm_index != d.csr_index) {
        if (a->tperm) {
            PICT_HIP_TRY(hipStreamSynchronize(s));
            PICT_HIP_TRY(hipFree(a->tperm));
            a->tperm = nullptr;
        }
        PICT_HIP_TRY(hipMalloc((void **)&a->tperm, sizeof(int32_t) * (size_t)(d.nnz > 0 ? d.nnz : 1)));
        pict_csr_desc C = csr_of(a, d.C_value);
        pict_status st = pict_csr_transpose_permutation(&C, a->tperm, stream);
        if (st) return st;
        a->tperm_row = d.csr_row;
        a->tperm_index = d.csr_index;
    }
    if (ticket) return run_async(a, 2, o, tape, velocity_grad, ticket, s);
    return run(a, 2, o, tape, velocity_grad, infos_out, s);
}

pict_status pict_piso_step_vjp(const pict_atlas *a, const pict_step_opts *o, const pict_step_tape *tape,
                               void *velocity_grad, pict_solve_info *infos_out, void *stream) {
    PICT_CHECK_ARG(infos_out, "NULL argument");
    return vjp_entry(a, o, tape, velocity_grad, infos_out, nullptr, stream);
}

pict_status pict_piso_step_vjp_async(const pict_atlas *a, const pict_step_opts *o, const pict_step_tape *tape,
                                     void *velocity_grad, int32_t *ticket, void *stream) {
    PICT_CHECK_ARG(ticket, "NULL argument");
    return vjp_entry(a, o, tape, velocity_grad, nullptr, ticket, stream);
}

pict_status pict_step_wait(int32_t ticket, pict_solve_info *infos_out) {
    PICT_CHECK_ARG(infos_out, "NULL argument");
    PICT_CHECK_ARG(ticket >= 0 && ticket < kPendingSteps, "no pending asynchronous step %d", ticket);
    PendingRing &ring = pending_ring();
    PendingStep &p = ring.slot[ticket];
    {
        std::lock_guard<std::mutex> lock(ring.mu);
        PICT_CHECK_ARG(p.busy, "no pending asynchronous step %d", ticket);
    }
    struct Release {
        PendingRing &r; PendingStep &p;
        ~Release() { std::lock_guard<std::mutex> lock(r.mu); p.busy = false; }
    } release{ring, p};
    PICT_HIP_TRY(hipEventSynchronize(p.ev));
    return collect(p.host, p.z, &p.o, p.D, infos_out, nullptr);
}

pict_status pict_subtract_mean(void *v, int32_t n, int32_t dtype, void *stream) {
    PICT_CHECK_ARG(v && n > 0, "vector missing");
    PICT_CHECK_ARG(dtype == PICT_DTYPE_F32 || dtype == PICT_DTYPE_F64, "unknown dtype %d", dtype);
    hipStream_t s = (hipStream_t)stream;
    double *part = nullptr;
    PICT_HIP_TRY(ws_alloc((void **)&part, sizeof(double) * kMeanBlocks, s));
    struct Guard {
        double *p; hipStream_t s;
        ~Guard() { if (p) (void)ws_free(p, s); }
    } guard{part, s};
    return dtype == PICT_DTYPE_F32 ? subtract_mean<float>(v, n, part, s) : subtract_mean<double>(v, n, part, s);
}

pict_status pict_set_step_timing(int32_t enable) {
    g_step_timing.enabled = enable != 0;
    return PICT_OK;
}

int32_t pict_last_step_solve_ms(double *ms_out, int32_t max_count) {
    const int n = (int)g_step_timing.ms.size();
    for (int i = 0; i < n && i < max_count; ++i) ms_out[i] = g_step_timing.ms[i];
    return n;
}

}  // extern "C"
