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

pict_status pict_piso_step(const pict_atlas *a, const pict_step_opts *o, pict_solve_info *infos_out, void *stream) {
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
    PICT_HIP_TRY(hipMallocAsync((void **)&dev, bytes, s));
    struct Guard {
        char *p; hipStream_t s;
        ~Guard() { if (p) (void)hipFreeAsync(p, s); }
    } guard{dev, s};
    double *part = (double *)(dev + slot * nsolves);
    hipEvent_t *ev = nullptr;
    pict_status st = timing_events(nsolves, ev);
    if (st) return st;

    const double dt = o->time_step;
    const pict_csr_desc C = csr_of(a, d.C_value), P = csr_of(a, d.P_value);
    // ---- advection (PISOtorch_simulation.py:1161-1188) ----
    if ((st = pict_setup_advection_matrix(a, dt, 0, stream))) return st;
    if ((st = pict_setup_advection_velocity(a, dt, 0, 0, stream))) return st;
    if ((st = f32 ? fill<float>(d.velocity_result, (size_t)D * N, 0.0, s) : fill<double>(d.velocity_result, (size_t)D * N, 0.0, s)))
        return st;
    if ((st = solve_enqueue(&C, d.velocity_rhs, d.velocity_result, D, &o->advection, dev,
                            (int *)(dev + 4 * sb), ev ? ev[0] : nullptr, ev ? ev[1] : nullptr, s)))
        return st;
    // ---- correctors (:1229-1320) ----
    for (int k = 0; k < K; ++k) {
        char *sl = dev + slot * (1 + k);
        if ((st = pict_setup_pressure_correction(a, dt, 0, 0, o->time_step_norm, stream))) return st;
        if ((st = f32 ? fill<float>(d.pressure_result, (size_t)N, 0.0, s) : fill<double>(d.pressure_result, (size_t)N, 0.0, s)))
            return st;
        if ((st = solve_enqueue(&P, d.pressure_rhs_div, d.pressure_result, 1, &o->pressure, sl, (int *)(sl + 4 * sb),
                                ev ? ev[2 * (1 + k)] : nullptr, ev ? ev[2 * (1 + k) + 1] : nullptr, s)))
            return st;
        if (o->normalize_pressure) {
            if ((st = f32 ? subtract_mean<float>(d.pressure_result, N, part, s)
                          : subtract_mean<double>(d.pressure_result, N, part, s)))
                return st;
        }
        if ((st = pict_copy_pressure_result_to_blocks(a, stream))) return st;
        if ((st = pict_correct_velocity(a, dt, o->velocity_corrector, o->time_step_norm, stream))) return st;
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

pict_status pict_subtract_mean(void *v, int32_t n, int32_t dtype, void *stream) {
    PICT_CHECK_ARG(v && n > 0, "vector missing");
    PICT_CHECK_ARG(dtype == PICT_DTYPE_F32 || dtype == PICT_DTYPE_F64, "unknown dtype %d", dtype);
    hipStream_t s = (hipStream_t)stream;
    double *part = nullptr;
    PICT_HIP_TRY(hipMallocAsync((void **)&part, sizeof(double) * kMeanBlocks, s));
    struct Guard {
        double *p; hipStream_t s;
        ~Guard() { if (p) (void)hipFreeAsync(p, s); }
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
