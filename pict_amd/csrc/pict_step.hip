// pict_step.hip -- a whole PISO time step queued natively on one stream.
//
// The orthogonal branch of the reference driver (PISOtorch_simulation.py:1054-1327:
// advection matrix and RHS, BiCGStab/CG velocity solve, `corrector_steps` x {pressure
// matrix, RHS, divergence, pressure solve, mean subtraction, copy to blocks, velocity
// correction}, copy to blocks) with every stage and solve queued back to back: the
// host waits once per step, to read the solvers' final states, instead of once per
// solve, and no interpreter runs between the stages.
#include <hip/hip_runtime.h>

#include <vector>

#include "pict_host.hpp"
#include "pict_solver.hpp"

namespace pict {

constexpr int kMeanThreads = 256;
constexpr int kMeanBlocks = 128;

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

extern "C" {

static pict_status copy_dev(void *dst, const void *src, size_t bytes, hipStream_t s) {
    PICT_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
    return PICT_OK;
}

pict_status pict_piso_step(const pict_atlas *a, const pict_step_opts *o, const pict_step_tape *tape,
                           pict_solve_info *infos_out, void *stream) {
    PICT_CHECK_ARG(a && o && infos_out, "NULL argument");
    const pict_domain_desc &d = a->desc;
    PICT_CHECK_ARG(d.C_value && d.P_value && d.A && d.velocity_rhs && d.velocity_result && d.pressure_rhs &&
                       d.pressure_rhs_div && d.pressure_result,
                   "domain tensors missing");
    PICT_CHECK_ARG(o->corrector_steps >= 0 && o->corrector_steps <= 64, "corrector_steps must be 0..64");
    PICT_CHECK_ARG(o->advection.transpose_a == 0 && o->pressure.transpose_a == 0, "transposed solves in a step");
    hipStream_t s = (hipStream_t)stream;
    const int D = d.num_dims, N = d.num_cells, K = o->corrector_steps;
    const bool f32 = d.dtype == PICT_DTYPE_F32;
    const size_t sb = solve_state_bytes();
    const int nsolves = 1 + K;
    // device: per solve kMaxRhs states + an error flag; mean partial sums
    const size_t slot = ((4 * sb + sizeof(int)) + 255) & ~size_t(255);
    const size_t bytes = slot * nsolves + sizeof(double) * kMeanBlocks;
    char *dev = nullptr;
    PICT_HIP_TRY(ws_alloc((void **)&dev, bytes, s));
    struct Guard {
        char *p; hipStream_t s;
        ~Guard() { if (p) (void)ws_free(p, s); }
    } guard{dev, s};
    double *part = (double *)(dev + slot * nsolves);
    hipEvent_t *ev = nullptr;
    pict_status st = timing_events(nsolves, ev);
    if (st) return st;

    const double dt = o->time_step;
    const size_t es = f32 ? sizeof(float) : sizeof(double);
    const size_t vecD = es * D * N, vec1 = es * N, nzb = es * d.nnz;
    if (tape)
        PICT_CHECK_ARG(tape->C_value && tape->A && tape->P_value && tape->velocity_star &&
                           (K == 0 || (tape->pressure_rhs && tape->pressure_raw && tape->velocity_result)),
                       "tape buffers missing");
    const pict_csr_desc C = csr_of(a, d.C_value), P = csr_of(a, d.P_value);
    // ---- advection (PISOtorch_simulation.py:1161-1188) ----
    if ((st = pict_setup_advection_matrix(a, dt, 0, stream))) return st;
    if ((st = pict_setup_advection_velocity(a, dt, 0, 0, stream))) return st;
    if (tape && ((st = copy_dev(tape->C_value, d.C_value, nzb, s)) || (st = copy_dev(tape->A, d.A, vec1, s))))
        return st;
    if ((st = f32 ? fill<float>(d.velocity_result, (size_t)D * N, 0.0, s) : fill<double>(d.velocity_result, (size_t)D * N, 0.0, s)))
        return st;
    if ((st = solve_enqueue(&C, d.velocity_rhs, d.velocity_result, D, &o->advection, dev,
                            (int *)(dev + 4 * sb), ev ? ev[0] : nullptr, ev ? ev[1] : nullptr, s)))
        return st;
    if (tape && (st = copy_dev(tape->velocity_star, d.velocity_result, vecD, s))) return st;
    // ---- correctors (:1229-1320) ----
    for (int k = 0; k < K; ++k) {
        char *sl = dev + slot * (1 + k);
        if ((st = pict_setup_pressure_correction(a, dt, 0, 0, o->time_step_norm, stream))) return st;
        if (tape && k == 0 && (st = copy_dev(tape->P_value, d.P_value, nzb, s))) return st;
        if (tape && (st = copy_dev((char *)tape->pressure_rhs + vecD * k, d.pressure_rhs, vecD, s))) return st;
        if ((st = f32 ? fill<float>(d.pressure_result, (size_t)N, 0.0, s) : fill<double>(d.pressure_result, (size_t)N, 0.0, s)))
            return st;
        if ((st = solve_enqueue(&P, d.pressure_rhs_div, d.pressure_result, 1, &o->pressure, sl, (int *)(sl + 4 * sb),
                                ev ? ev[2 * (1 + k)] : nullptr, ev ? ev[2 * (1 + k) + 1] : nullptr, s)))
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

    // ---- one wait: the solvers' final states ----
    HostStates &h = g_host_states;
    if (h.bytes < slot * nsolves) {
        if (h.p) PICT_HIP_TRY(hipHostFree(h.p));
        h.p = nullptr;
        PICT_HIP_TRY(hipHostMalloc(&h.p, slot * nsolves, hipHostMallocDefault));
        h.bytes = slot * nsolves;
    }
    PICT_HIP_TRY(hipMemcpyAsync(h.p, dev, slot * nsolves, hipMemcpyDeviceToHost, s));
    PICT_HIP_TRY(hipStreamSynchronize(s));
    const char *hp = (const char *)h.p;
    int err = 0;
    for (int i = 0; i < nsolves; ++i) err |= *(const int *)(hp + slot * i + 4 * sb);
    if (err) {
        set_error("solver grid barrier exceeded its spin limit");
        return PICT_ERR_DEVICE_SYNC;
    }
    solve_state_to_info(hp, D, infos_out);
    for (int k = 0; k < K; ++k) solve_state_to_info(hp + slot * (1 + k), 1, infos_out + D + k);
    if (ev) {
        g_step_timing.ms.assign(nsolves, 0.0);
        for (int i = 0; i < nsolves; ++i) {
            float ms = 0.f;
            PICT_HIP_TRY(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
            g_step_timing.ms[i] = ms;
        }
    }
    return PICT_OK;
}

// Reverse pass of one pict_piso_step (the adjoint of the orthogonal step; the order in which
// PISOtorch_diff's autograd graph visits the stages, PISOtorch_diff.py:295-1234).
pict_status pict_piso_step_vjp(const pict_atlas *a, const pict_step_opts *o, const pict_step_tape *tape,
                               void *velocity_grad, pict_solve_info *infos_out, void *stream) {
    PICT_CHECK_ARG(a && o && tape && velocity_grad && infos_out, "NULL argument");
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
    const int D = d.num_dims, N = d.num_cells, K = o->corrector_steps;
    const bool f32 = d.dtype == PICT_DTYPE_F32;
    const size_t es = f32 ? sizeof(float) : sizeof(double);
    const size_t vecD = es * D * N, vec1 = es * N, nzb = es * d.nnz;
    const size_t sb = solve_state_bytes();
    const int nsolves = 1 + K;
    const size_t slot = ((4 * sb + sizeof(int)) + 255) & ~size_t(255);
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    // workspace: states | mean partials | g_u [D*N] | A_tot [N] | C_tot [nnz] | lambda [D*N]
    const size_t o_part = slot * nsolves, o_gu = o_part + al(sizeof(double) * kMeanBlocks), o_at = o_gu + al(vecD),
                 o_ct = o_at + al(vec1), o_lam = o_ct + al(nzb), bytes = o_lam + al(vecD);
    char *dev = nullptr;
    PICT_HIP_TRY(ws_alloc((void **)&dev, bytes, s));
    struct Guard {
        char *p; hipStream_t s;
        ~Guard() { if (p) (void)ws_free(p, s); }
    } guard{dev, s};
    double *part = (double *)(dev + o_part);
    void *g_u = dev + o_gu, *A_tot = dev + o_at, *C_tot = dev + o_ct, *lam = dev + o_lam;
    const double dt = o->time_step;
    hipEvent_t *ev = nullptr;
    pict_status st = timing_events(nsolves, ev);
    if (st) return st;
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
    // transposed matrices share the pattern's A^T map, built on the atlas's first reverse pass
    pict_csr_desc C = csr_of(a, d.C_value), P = csr_of(a, d.P_value);
    if (!a->tperm || a->tperm_row != d.csr_row || a->tperm_index != d.csr_index) {
        if (a->tperm) {
            PICT_HIP_TRY(hipStreamSynchronize(s));
            PICT_HIP_TRY(hipFree(a->tperm));
            a->tperm = nullptr;
        }
        PICT_HIP_TRY(hipMalloc((void **)&a->tperm, sizeof(int32_t) * (size_t)(d.nnz > 0 ? d.nnz : 1)));
        PICT_TRY(pict_csr_transpose_permutation(&C, a->tperm, stream));
        a->tperm_row = d.csr_row;
        a->tperm_index = d.csr_index;
    }
    C.transpose_perm = a->tperm;
    P.transpose_perm = a->tperm;
    pict_solve_opts adv = o->advection, prs = o->pressure;
    adv.transpose_a = 1;
    prs.transpose_a = 1;
    adv.zero_rhs_shortcut = 1;
    prs.zero_rhs_shortcut = 1;
    prs.return_best_result = o->pressure.return_best_result;

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
        char *sl = dev + slot * (1 + k);
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
                               ev ? ev[2 * (1 + k) + 1] : nullptr, s));
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
                           ev ? ev[0] : nullptr, ev ? ev[1] : nullptr, s));
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

    // one wait: the adjoint solvers' final states
    HostStates &h = g_host_states;
    if (h.bytes < slot * nsolves) {
        if (h.p) PICT_HIP_TRY(hipHostFree(h.p));
        h.p = nullptr;
        PICT_HIP_TRY(hipHostMalloc(&h.p, slot * nsolves, hipHostMallocDefault));
        h.bytes = slot * nsolves;
    }
    PICT_HIP_TRY(hipMemcpyAsync(h.p, dev, slot * nsolves, hipMemcpyDeviceToHost, s));
    PICT_HIP_TRY(hipStreamSynchronize(s));
    const char *hp = (const char *)h.p;
    int err = 0;
    for (int i = 0; i < nsolves; ++i) err |= *(const int *)(hp + slot * i + 4 * sb);
    if (err) {
        set_error("solver grid barrier exceeded its spin limit");
        return PICT_ERR_DEVICE_SYNC;
    }
    solve_state_to_info(hp, D, infos_out);
    for (int k = 0; k < K; ++k) solve_state_to_info(hp + slot * (1 + k), 1, infos_out + D + k);
    if (ev) {
        g_step_timing.ms.assign(nsolves, 0.0);
        for (int i = 0; i < nsolves; ++i) {
            float ms = 0.f;
            PICT_HIP_TRY(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
            g_step_timing.ms[i] = ms;
        }
    }
    return PICT_OK;
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
