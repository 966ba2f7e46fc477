// pict_batch.hip -- many independent PISO simulations stepped together.
//
// A batch holds `count` domains of one dtype and dimension (any block structure,
// size, viscosity and boundaries).  One forward step of all of them runs
// pict_piso_step's sequence with
//   - one launch per stage over the tiles of every domain: each workgroup reads its
//     domain's view and time step from a table, then runs the single-domain cell
//     body (pict_cells.hpp) on one or more 256-cell tiles; stages that only read
//     data written by earlier launches share a launch;
//   - one cooperative launch per solve (pict_solve.hip, batch_solve_*): every
//     resident workgroup runs its list of (system, workgroup) jobs;
//   - per-domain mean subtraction in the single-domain step's summation order.
// The step is captured into a HIP graph on its second call and replayed; its tables
// live in device memory and are rewritten in place when a domain's tensors change,
// so the graph stays valid.  The host waits once per step, for the solver states.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "pict_cells.hpp"
#include "pict_host.hpp"
#include "pict_solver.hpp"

namespace pict {

// one domain of the batch as the kernels see it
template <typename R>
struct BatchDom {
    DomainK<R> k;
    R dt;
};

// One workgroup per `tpw` consecutive entries of the tile list (domain, tile of that domain).
template <typename R, class Op>
__global__ __launch_bounds__(kCellsPerTile) void kb_cells(const BatchDom<R> *__restrict__ doms,
                                                          const int2 *__restrict__ tiles, int ntiles, int tpw, Op op) {
    for (int t = 0; t < tpw; ++t) {
        const int j = (int)blockIdx.x * tpw + t;
        if (j >= ntiles) return;
        const int2 tl = tiles[j];
        const BatchDom<R> d = doms[tl.x];
        op(d, tl.y);
    }
}

// SetupAdvectionMatrix + SetupAdvectionVelocity + zero start vector (independent per cell)
template <typename R, int D>
struct OpAdvection {
    __device__ void operator()(const BatchDom<R> &d, int t) const {
        adv_matrix_cell<R, D>(d.k, t, d.dt);
        adv_rhs_cell<R, D>(d.k, t, d.dt, 0);
        zero_result_cell<R>(d.k, t, 0);
    }
};
// pressure matrix (it depends on A only: built for the first corrector) + pressure RHS + zero start vector
template <typename R, int D>
struct OpPressureSetup {
    int matrix;
    __device__ void operator()(const BatchDom<R> &d, int t) const {
        if (matrix) p_matrix_cell<R, D>(d.k, t);
        p_rhs_cell<R, D>(d.k, t, d.dt);
        zero_result_cell<R>(d.k, t, 1);
    }
};
template <typename R, int D>
struct OpDivergence {
    int tsn;
    __device__ void operator()(const BatchDom<R> &d, int t) const {
        divergence_cell<R, D>(d.k, t, d.k.prhs, d.k.prhs_div, d.dt, tsn);
    }
};
template <typename R>
struct OpCopyToBlocks {
    int field;  // 0 velocity, 1 pressure
    __device__ void operator()(const BatchDom<R> &d, int t) const {
        copy_result_cell<R>(d.k, t, field == 0 ? d.k.vres : d.k.pres, field, 1);
    }
};
template <typename R, int D>
struct OpCorrect {
    int tsn;
    __device__ void operator()(const BatchDom<R> &d, int t) const { correct_velocity_cell<R, D>(d.k, t, d.dt, tsn); }
};

// Per-domain FP64 partial sums of the pressure result: grid (kMeanBlocks, count), the single-domain
// step's shape and order for each domain.
template <typename R>
__global__ __launch_bounds__(kMeanThreads) void kb_partial_sum(const BatchDom<R> *__restrict__ doms, double *part) {
    __shared__ double lds[kMeanThreads / 64];
    const R *v = doms[blockIdx.y].k.pres;
    const int n = doms[blockIdx.y].k.N;
    double s = 0.0;
    for (int i = blockIdx.x * kMeanThreads + threadIdx.x; i < n; i += gridDim.x * kMeanThreads) s += (double)v[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kMeanThreads / 64; ++w) t += lds[w];
        part[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = t;
    }
}
template <typename R>
__global__ __launch_bounds__(kMeanThreads) void kb_sub_mean(const BatchDom<R> *__restrict__ doms, const double *part) {
    __shared__ double mean;
    R *v = doms[blockIdx.y].k.pres;
    const int n = doms[blockIdx.y].k.N;
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (unsigned b = 0; b < gridDim.x; ++b) t += part[(size_t)blockIdx.y * gridDim.x + b];
        mean = t / (double)n;
    }
    __syncthreads();
    const R m = (R)mean;
    for (int i = blockIdx.x * kMeanThreads + threadIdx.x; i < n; i += gridDim.x * kMeanThreads) v[i] -= m;
}

}  // namespace pict

using namespace pict;

// The plan of a batch for one set of step options: its solves, the device slots of their
// states and error flags, the mean partial sums, and the captured graph.
struct BatchPlan {
    pict_step_opts opts;
    pict_batch_config cfg;
    BatchSolve *adv = nullptr, *prs = nullptr;
    char *dev = nullptr;    // [advection states | pressure states per corrector | error flags | mean partials]
    void *host = nullptr;   // pinned copy of [states | error flags]
    size_t o_prs = 0, o_err = 0, o_part = 0, host_bytes = 0, dev_bytes = 0;
    hipGraphExec_t exec = nullptr;
    int steps = 0;          // steps run with this plan
    bool rescheduled = false;
    ~BatchPlan() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (adv) batch_solve_destroy(adv);
        if (prs) batch_solve_destroy(prs);
        if (dev) (void)hipFree(dev);
        if (host) (void)hipHostFree(host);
    }
};

struct pict_batch {
    int count = 0, D = 0, dtype = 0, device = 0;
    std::vector<const pict_atlas *> atlases;
    int ntiles = 0;
    char *dev = nullptr;            // [domain table | tile list]
    size_t dom_stride = 0, o_tiles = 0;
    std::vector<char> dom_host;     // the uploaded domain table
    bool uploaded = false;
    pict_batch_config cfg{1, 0, 0, 1, 0, 0};
    std::unique_ptr<BatchPlan> plan;
    hipStream_t capture = nullptr;
    bool last_graph = false;
    // diagnostics: HIP events around every solve launch (steps then run without the graph)
    bool timing = false;
    std::vector<hipEvent_t> ev;
    std::vector<double> solve_ms;
    ~pict_batch() {
        plan.reset();
        if (dev) (void)hipFree(dev);
        if (capture) (void)hipStreamDestroy(capture);
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
};

namespace {

bool batch_graphs_enabled() {
    const char *e = getenv("PICT_STEP_GRAPHS");
    return !e || atoi(e) != 0;
}

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

pict_status check_domain(const pict_atlas *a, int s) {
    const pict_domain_desc &d = a->desc;
    PICT_CHECK_ARG(d.C_value && d.P_value && d.A && d.velocity_rhs && d.velocity_result && d.pressure_rhs &&
                       d.pressure_rhs_div && d.pressure_result && d.csr_row && d.csr_index,
                   "batch domain %d: domain tensors missing", s);
    return PICT_OK;
}

template <typename R>
void build_table(const pict_batch *b, const double *dts, double dt0, std::vector<char> &out) {
    out.assign(b->dom_stride * b->count, 0);
    for (int s = 0; s < b->count; ++s) {
        BatchDom<R> d;
        memset(&d, 0, sizeof d);
        d.k = b->atlases[s]->view<R>();
        d.dt = (R)(dts ? dts[s] : dt0);
        memcpy(out.data() + b->dom_stride * s, &d, sizeof d);
    }
}

// the domain table from the atlases' current descriptors and the time steps; rewritten (after the
// queued work that reads it) when it changed
pict_status refresh_tables(pict_batch *b, const double *dts, double dt0, hipStream_t s) {
    std::vector<char> t;
    if (b->dtype == PICT_DTYPE_F32) build_table<float>(b, dts, dt0, t);
    else build_table<double>(b, dts, dt0, t);
    if (b->uploaded && t == b->dom_host) return PICT_OK;
    PICT_HIP_TRY(hipStreamSynchronize(s));
    PICT_HIP_TRY(hipMemcpyAsync(b->dev, t.data(), t.size(), hipMemcpyHostToDevice, s));
    PICT_HIP_TRY(hipStreamSynchronize(s));
    b->dom_host.swap(t);
    b->uploaded = true;
    return PICT_OK;
}

// the solve vectors of every domain
struct SolveSet {
    std::vector<pict_csr_desc> A;
    std::vector<const void *> rhs;
    std::vector<void *> x;
};
SolveSet advection_set(const pict_batch *b) {
    SolveSet v;
    for (const pict_atlas *a : b->atlases) {
        v.A.push_back(csr_of(a, a->desc.C_value));
        v.rhs.push_back(a->desc.velocity_rhs);
        v.x.push_back(a->desc.velocity_result);
    }
    return v;
}
SolveSet pressure_set(const pict_batch *b) {
    SolveSet v;
    for (const pict_atlas *a : b->atlases) {
        v.A.push_back(csr_of(a, a->desc.P_value));
        v.rhs.push_back(a->desc.pressure_rhs_div);
        v.x.push_back(a->desc.pressure_result);
    }
    return v;
}

pict_status create_solve(const pict_batch *b, bool pressure, const pict_step_opts *o, const pict_batch_config &cfg,
                         BatchSolve **out, hipStream_t s) {
    const SolveSet v = pressure ? pressure_set(b) : advection_set(b);
    BatchSolveConfig c;
    c.rows_per_thread = cfg.pressure_rows_per_thread;
    c.rows_per_wg = cfg.advection_rows_per_workgroup;
    // BiCGStab in registers (where the kernel exists for the system; else the streaming kernel)
    const int adv = cfg.advection_rows_per_thread == 0 ? 2 : cfg.advection_rows_per_thread;
    if (!pressure) c.bicg_rows_per_thread = adv > 0 ? adv : 0;
    else if (o->pressure.use_bicg)
        c.bicg_rows_per_thread = cfg.pressure_rows_per_thread == 0 ? 2
                                 : cfg.pressure_rows_per_thread <= 4 ? cfg.pressure_rows_per_thread : 0;
    return batch_solve_create(v.A.data(), v.rhs.data(), v.x.data(), b->count, pressure ? 1 : b->D,
                              pressure ? &o->pressure : &o->advection, c, nullptr, out, s);
}

pict_status update_solves(pict_batch *b, hipStream_t s) {
    BatchPlan &p = *b->plan;
    const SolveSet va = advection_set(b), vp = pressure_set(b);
    pict_status st = batch_solve_update(p.adv, va.A.data(), va.rhs.data(), va.x.data(), s);
    if (st) return st;
    return batch_solve_update(p.prs, vp.A.data(), vp.rhs.data(), vp.x.data(), s);
}

bool same_opts(const pict_step_opts &x, const pict_step_opts &y) {
    // the time step lives in the domain table, not in the plan
    pict_step_opts a = x, c = y;
    a.time_step = c.time_step = 0.0;
    return memcmp(&a, &c, sizeof a) == 0;
}

pict_status ensure_plan(pict_batch *b, const pict_step_opts *o, hipStream_t s) {
    if (b->plan && same_opts(b->plan->opts, *o) && memcmp(&b->plan->cfg, &b->cfg, sizeof b->cfg) == 0)
        return update_solves(b, s);
    PICT_HIP_TRY(hipStreamSynchronize(s));
    b->plan.reset();
    std::unique_ptr<BatchPlan> p(new BatchPlan());
    p->opts = *o;
    p->cfg = b->cfg;
    pict_status st = create_solve(b, false, o, b->cfg, &p->adv, s);
    if (st) return st;
    if ((st = create_solve(b, true, o, b->cfg, &p->prs, s))) return st;
    auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
    const size_t sb = solve_state_bytes();
    const int K = o->corrector_steps;
    p->o_prs = al(sb * (size_t)b->count * b->D);
    p->o_err = p->o_prs + al(sb * (size_t)b->count * (K > 0 ? K : 1));
    p->host_bytes = p->o_err + sizeof(int) * (1 + K);
    p->o_part = al(p->host_bytes);
    p->dev_bytes = p->o_part + sizeof(double) * (size_t)kMeanBlocks * b->count;
    PICT_HIP_TRY(hipMalloc((void **)&p->dev, p->dev_bytes));
    PICT_HIP_TRY(hipHostMalloc(&p->host, p->host_bytes, hipHostMallocDefault));
    b->plan = std::move(p);
    return PICT_OK;
}

template <typename R, class Op>
pict_status cells(const pict_batch *b, int tpw, Op op, hipStream_t s) {
    const unsigned grid = (unsigned)((b->ntiles + tpw - 1) / tpw);
    hipLaunchKernelGGL((kb_cells<R, Op>), dim3(grid), dim3(kCellsPerTile), 0, s, (const BatchDom<R> *)b->dev,
                       (const int2 *)(b->dev + b->o_tiles), b->ntiles, tpw, op);
    return check_launch();
}

// The stage launches of a step (no solve): `part` 0 advection setup, 1 pressure setup of corrector
// k (matrix when k == 0), 2 divergence, 3 mean subtraction, 4 pressure to blocks + velocity
// correction, 5 velocity to blocks, 6 velocity correction alone (the autotuner's, which leaves
// the block tensors alone).
template <typename R, int D>
pict_status stage(const pict_batch *b, const pict_step_opts *o, int part, int k, int tpw, hipStream_t s) {
    switch (part) {
        case 0: return cells<R>(b, tpw, OpAdvection<R, D>{}, s);
        case 1: return cells<R>(b, tpw, OpPressureSetup<R, D>{k == 0 ? 1 : 0}, s);
        case 2: return cells<R>(b, tpw, OpDivergence<R, D>{o->time_step_norm}, s);
        case 3: {
            double *part_sums = (double *)(b->plan->dev + b->plan->o_part);
            hipLaunchKernelGGL(kb_partial_sum<R>, dim3(kMeanBlocks, b->count), dim3(kMeanThreads), 0, s,
                               (const BatchDom<R> *)b->dev, part_sums);
            hipLaunchKernelGGL(kb_sub_mean<R>, dim3(kMeanBlocks, b->count), dim3(kMeanThreads), 0, s,
                               (const BatchDom<R> *)b->dev, (const double *)part_sums);
            return check_launch();
        }
        case 4: {
            pict_status st = cells<R>(b, tpw, OpCopyToBlocks<R>{1}, s);
            if (st) return st;
            return cells<R>(b, tpw, OpCorrect<R, D>{o->time_step_norm}, s);
        }
        case 5: return cells<R>(b, tpw, OpCopyToBlocks<R>{0}, s);
        default: return cells<R>(b, tpw, OpCorrect<R, D>{o->time_step_norm}, s);
    }
}

pict_status run_stage(const pict_batch *b, const pict_step_opts *o, int part, int k, int tpw, hipStream_t s) {
    const bool f32 = b->dtype == PICT_DTYPE_F32;
    switch (b->D) {
        case 1: return f32 ? stage<float, 1>(b, o, part, k, tpw, s) : stage<double, 1>(b, o, part, k, tpw, s);
        case 2: return f32 ? stage<float, 2>(b, o, part, k, tpw, s) : stage<double, 2>(b, o, part, k, tpw, s);
        default: return f32 ? stage<float, 3>(b, o, part, k, tpw, s) : stage<double, 3>(b, o, part, k, tpw, s);
    }
}

// the whole step queued on s (allocates nothing, never waits: capturable); ev (may be null): 2 events per
// solve, recorded around its launch
pict_status enqueue_step(const pict_batch *b, const pict_step_opts *o, hipStream_t s, hipEvent_t *ev = nullptr) {
    const BatchPlan &p = *b->plan;
    const int tpw = p.cfg.tiles_per_workgroup > 0 ? p.cfg.tiles_per_workgroup : 1;
    const size_t sb = solve_state_bytes();
    const int K = o->corrector_steps;
    int *err = (int *)(p.dev + p.o_err);
    pict_status st;
#define PICT_TRY(expr) do { if ((st = (expr))) return st; } while (0)
    PICT_TRY(run_stage(b, o, 0, 0, tpw, s));
    PICT_TRY(batch_solve_enqueue(p.adv, p.dev, err, ev ? ev[0] : nullptr, ev ? ev[1] : nullptr, s));
    for (int k = 0; k < K; ++k) {
        PICT_TRY(run_stage(b, o, 1, k, tpw, s));
        PICT_TRY(run_stage(b, o, 2, k, tpw, s));
        PICT_TRY(batch_solve_enqueue(p.prs, p.dev + p.o_prs + sb * (size_t)b->count * k, err + 1 + k,
                                     ev ? ev[2 * (1 + k)] : nullptr, ev ? ev[2 * (1 + k) + 1] : nullptr, s));
        if (o->normalize_pressure) PICT_TRY(run_stage(b, o, 3, k, tpw, s));
        PICT_TRY(run_stage(b, o, 4, k, tpw, s));
    }
    PICT_TRY(run_stage(b, o, 5, 0, tpw, s));
#undef PICT_TRY
    PICT_HIP_TRY(hipMemcpyAsync(p.host, p.dev, p.host_bytes, hipMemcpyDeviceToHost, s));
    return PICT_OK;
}

pict_status collect(pict_batch *b, const pict_step_opts *o, pict_solve_info *infos, hipStream_t s) {
    BatchPlan &p = *b->plan;
    const char *h = (const char *)p.host;
    const int K = o->corrector_steps, D = b->D, per = D + K;
    const size_t sb = solve_state_bytes();
    const int *err = (const int *)(h + p.o_err);
    for (int i = 0; i <= K; ++i)
        if (err[i]) {
            set_error("batched solver barrier exceeded its spin limit (solve %d)", i);
            return PICT_ERR_DEVICE_SYNC;
        }
    for (int d = 0; d < b->count; ++d) {
        solve_state_to_info(h + sb * (size_t)d * D, D, infos + (size_t)d * per);
        for (int k = 0; k < K; ++k)
            solve_state_to_info(h + p.o_prs + sb * ((size_t)b->count * k + d), 1, infos + (size_t)d * per + D + k);
    }
    // after the first step: schedule the solves by the iterations they took
    if (p.cfg.adaptive_schedule && !p.rescheduled) {
        std::vector<double> ca(b->count), cp(b->count);
        for (int d = 0; d < b->count; ++d) {
            int ia = 1, ip = 1;
            for (int c = 0; c < D; ++c) ia = std::max(ia, infos[(size_t)d * per + c].used_iterations + 1);
            for (int k = 0; k < K; ++k) ip = std::max(ip, infos[(size_t)d * per + D + k].used_iterations + 1);
            const int na = b->atlases[d]->desc.num_cells;
            ca[d] = (double)ia * (2048.0 + (double)na / std::max(1, (na + 1023) / 1024));
            cp[d] = (double)ip;
        }
        pict_status st = batch_solve_reschedule(p.adv, ca.data(), s);
        if (st) return st;
        if (K > 0 && (st = batch_solve_reschedule(p.prs, cp.data(), s))) return st;
        p.rescheduled = true;
    }
    return PICT_OK;
}

pict_status capture_step(pict_batch *b, const pict_step_opts *o) {
    BatchPlan &p = *b->plan;
    if (!b->capture) PICT_HIP_TRY(hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking));
    hipGraph_t graph = nullptr;
    PICT_HIP_TRY(hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal));
    pict_status st = enqueue_step(b, o, b->capture);
    const hipError_t ce = hipStreamEndCapture(b->capture, &graph);
    if (st || ce != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        if (st) return st;
        set_error("batch step capture failed: %s", hipGetErrorString(ce));
        return PICT_ERR_HIP;
    }
    const hipError_t ie = hipGraphInstantiate(&p.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
        p.exec = nullptr;
        set_error("batch step graph instantiation failed: %s", hipGetErrorString(ie));
        return PICT_ERR_HIP;
    }
    return PICT_OK;
}

pict_status check_opts(const pict_batch *b, const pict_step_opts *o) {
    PICT_CHECK_ARG(o->corrector_steps >= 0 && o->corrector_steps <= 64, "corrector_steps must be 0..64");
    PICT_CHECK_ARG(o->advection.transpose_a == 0 && o->pressure.transpose_a == 0, "transposed solves in a step");
    PICT_CHECK_ARG(o->velocity_corrector == 0 || o->velocity_corrector == 1, "velocity corrector must be 0 or 1");
    for (int s = 0; s < b->count; ++s) {
        pict_status st = check_domain(b->atlases[s], s);
        if (st) return st;
    }
    return PICT_OK;
}

float elapsed(hipEvent_t e0, hipEvent_t e1) {
    float ms = 0.f;
    return hipEventElapsedTime(&ms, e0, e1) == hipSuccess ? ms : 1e30f;
}

}  // namespace

extern "C" {

pict_status pict_batch_create(const pict_atlas *const *atlases, int32_t count, pict_batch **out) {
    PICT_CHECK_ARG(atlases && out, "NULL argument");
    PICT_CHECK_ARG(count >= 1 && count <= 65535, "batch size must be 1..65535 (got %d)", count);
    std::unique_ptr<pict_batch> b(new pict_batch());
    b->count = count;
    for (int s = 0; s < count; ++s) {
        const pict_atlas *a = atlases[s];
        PICT_CHECK_ARG(a != nullptr, "atlas %d is NULL", s);
        PICT_CHECK_ARG(a->desc.dtype == atlases[0]->desc.dtype && a->desc.num_dims == atlases[0]->desc.num_dims &&
                           a->device == atlases[0]->device,
                       "atlas %d: dtype, dimension or device differs from atlas 0", s);
        b->atlases.push_back(a);
    }
    b->D = atlases[0]->desc.num_dims;
    b->dtype = atlases[0]->desc.dtype;
    b->device = atlases[0]->device;
    std::vector<int2> tiles;
    for (int s = 0; s < count; ++s)
        for (int t = 0; t < atlases[s]->num_tiles; ++t) tiles.push_back(make_int2(s, t));
    b->ntiles = (int)tiles.size();
    auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
    b->dom_stride = b->dtype == PICT_DTYPE_F32 ? sizeof(BatchDom<float>) : sizeof(BatchDom<double>);
    b->o_tiles = al(b->dom_stride * count);
    const size_t bytes = b->o_tiles + sizeof(int2) * tiles.size();
    PICT_HIP_TRY(hipMalloc((void **)&b->dev, bytes));
    PICT_HIP_TRY(hipMemcpy(b->dev + b->o_tiles, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice));
    *out = b.release();
    return PICT_OK;
}

pict_status pict_batch_destroy(pict_batch *b) {
    if (!b) return PICT_OK;
    (void)hipDeviceSynchronize();
    delete b;
    return PICT_OK;
}

pict_status pict_batch_set_config(pict_batch *b, const pict_batch_config *c) {
    PICT_CHECK_ARG(b && c, "NULL argument");
    PICT_CHECK_ARG(c->tiles_per_workgroup >= 0 && c->tiles_per_workgroup <= 16, "tiles_per_workgroup must be 0..16");
    PICT_CHECK_ARG(c->pressure_rows_per_thread == 0 || c->pressure_rows_per_thread == 1 ||
                       c->pressure_rows_per_thread == 2 || c->pressure_rows_per_thread == 4 ||
                       c->pressure_rows_per_thread == 8,
                   "pressure_rows_per_thread must be 0, 1, 2, 4 or 8");
    PICT_CHECK_ARG(c->advection_rows_per_workgroup == 0 || c->advection_rows_per_workgroup >= 64,
                   "advection_rows_per_workgroup must be 0 or >= 64");
    PICT_CHECK_ARG(c->advection_rows_per_thread >= -1 && c->advection_rows_per_thread <= 4 &&
                       c->advection_rows_per_thread != 3,
                   "advection_rows_per_thread must be -1, 0, 1, 2 or 4");
    b->cfg = *c;
    return PICT_OK;
}

pict_status pict_batch_get_config(const pict_batch *b, pict_batch_config *c) {
    PICT_CHECK_ARG(b && c, "NULL argument");
    *c = b->cfg;
    return PICT_OK;
}

pict_status pict_batch_step(pict_batch *b, const pict_step_opts *o, const double *time_steps,
                            pict_solve_info *infos_out, void *stream) {
    PICT_CHECK_ARG(b && o && infos_out, "NULL argument");
    hipStream_t s = (hipStream_t)stream;
    pict_status st = check_opts(b, o);
    if (st) return st;
    if ((st = refresh_tables(b, time_steps, o->time_step, s))) return st;
    if ((st = ensure_plan(b, o, s))) return st;
    BatchPlan &p = *b->plan;
    b->last_graph = false;
    const int nsolves = 1 + o->corrector_steps;
    if (b->timing) {
        while ((int)b->ev.size() < 2 * nsolves) {
            hipEvent_t e;
            PICT_HIP_TRY(hipEventCreate(&e));
            b->ev.push_back(e);
        }
        if ((st = enqueue_step(b, o, s, b->ev.data()))) return st;
    } else if (batch_graphs_enabled() && p.steps >= 1) {
        if (!p.exec && (st = capture_step(b, o))) return st;
        PICT_HIP_TRY(hipGraphLaunch(p.exec, s));
        b->last_graph = true;
    } else {
        if ((st = enqueue_step(b, o, s))) return st;
    }
    PICT_HIP_TRY(hipStreamSynchronize(s));
    ++p.steps;
    if (b->timing) {
        b->solve_ms.assign(nsolves, 0.0);
        for (int i = 0; i < nsolves; ++i) b->solve_ms[i] = elapsed(b->ev[2 * i], b->ev[2 * i + 1]);
    }
    return collect(b, o, infos_out, s);
}

pict_status pict_batch_tune(pict_batch *b, const pict_step_opts *o, const double *time_steps, int32_t repeats,
                            pict_batch_config *chosen, void *stream) {
    PICT_CHECK_ARG(b && o, "NULL argument");
    PICT_CHECK_ARG(repeats >= 1 && repeats <= 100, "repeats must be 1..100");
    hipStream_t s = (hipStream_t)stream;
    pict_status st = check_opts(b, o);
    if (st) return st;
    if ((st = refresh_tables(b, time_steps, o->time_step, s))) return st;
    if ((st = ensure_plan(b, o, s))) return st;  // the mean partials and the default solves
    hipEvent_t e0, e1;
    PICT_HIP_TRY(hipEventCreate(&e0));
    PICT_HIP_TRY(hipEventCreate(&e1));
    struct Ev {
        hipEvent_t a, b;
        ~Ev() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } ev{e0, e1};
    pict_batch_config best = b->cfg;
    auto timed = [&](auto &&fn, float &ms) -> pict_status {
        ms = 1e30f;
        for (int r = 0; r < repeats; ++r) {
            PICT_HIP_TRY(hipEventRecord(e0, s));
            pict_status q = fn();
            if (q) return q;
            PICT_HIP_TRY(hipEventRecord(e1, s));
            PICT_HIP_TRY(hipEventSynchronize(e1));
            ms = std::min(ms, elapsed(e0, e1));
        }
        return PICT_OK;
    };
    // 1. stage kernels: tiles per workgroup (the advection and pressure setups, divergence, correction)
    {
        float bt = 1e30f;
        for (int tpw : {1, 2, 4}) {
            float ms;
            st = timed([&]() -> pict_status {
                pict_status q;
                if ((q = run_stage(b, o, 0, 0, tpw, s))) return q;
                if ((q = run_stage(b, o, 1, 0, tpw, s))) return q;
                if ((q = run_stage(b, o, 2, 0, tpw, s))) return q;
                return run_stage(b, o, 6, 0, tpw, s);
            }, ms);
            if (st) return st;
            if (ms < bt) {
                bt = ms;
                best.tiles_per_workgroup = tpw;
            }
        }
    }
    // 2. advection BiCGStab: rows per thread of the register-resident kernel, rows per workgroup of the
    // streaming one
    {
        float bt = 1e30f;
        const int cand[][2] = {{1, 0}, {2, 0}, {4, 0}, {-1, 512}, {-1, 1024}, {-1, 2048}};
        for (const auto &cv : cand) {
            pict_batch_config c = best;
            c.advection_rows_per_thread = cv[0];
            c.advection_rows_per_workgroup = cv[1];
            BatchSolve *bs = nullptr;
            st = create_solve(b, false, o, c, &bs, s);
            if (st == PICT_ERR_UNSUPPORTED || st == PICT_ERR_INVALID) continue;  // a system too large for it
            if (st) return st;
            float ms;
            st = timed([&]() -> pict_status {
                pict_status q = run_stage(b, o, 0, 0, best.tiles_per_workgroup, s);
                if (q) return q;
                return batch_solve_enqueue(bs, nullptr, nullptr, nullptr, nullptr, s);
            }, ms);
            batch_solve_destroy(bs);
            if (st) return st;
            if (ms < bt) {
                bt = ms;
                best.advection_rows_per_thread = cv[0];
                best.advection_rows_per_workgroup = cv[1];
            }
        }
    }
    // 3. pressure CG: rows per thread (a BiCGStab pressure solve takes the pressure rows per thread as its
    // register-resident shape, else the advection's rows per workgroup)
    if (o->corrector_steps > 0 && !o->pressure.use_bicg) {
        float bt = 1e30f;
        for (int v : {1, 2, 4, 8}) {
            pict_batch_config c = best;
            c.pressure_rows_per_thread = v;
            BatchSolve *bs = nullptr;
            st = create_solve(b, true, o, c, &bs, s);
            if (st == PICT_ERR_UNSUPPORTED || st == PICT_ERR_INVALID) continue;  // over the register budget or too big
            if (st) return st;
            float ms;
            st = timed([&]() -> pict_status {
                pict_status q = run_stage(b, o, 1, 0, best.tiles_per_workgroup, s);
                if (q) return q;
                if ((q = run_stage(b, o, 2, 0, best.tiles_per_workgroup, s))) return q;
                return batch_solve_enqueue(bs, nullptr, nullptr, nullptr, nullptr, s);
            }, ms);
            batch_solve_destroy(bs);
            if (st) return st;
            if (ms < bt) {
                bt = ms;
                best.pressure_rows_per_thread = v;
            }
        }
    }
    PICT_HIP_TRY(hipStreamSynchronize(s));
    b->cfg = best;
    b->plan.reset();
    if (chosen) *chosen = best;
    return PICT_OK;
}

pict_status pict_batch_set_timing(pict_batch *b, int32_t enable) {
    PICT_CHECK_ARG(b, "NULL argument");
    b->timing = enable != 0;
    return PICT_OK;
}

int32_t pict_batch_last_solve_ms(const pict_batch *b, double *ms_out, int32_t max_count) {
    if (!b) return 0;
    const int n = (int)b->solve_ms.size();
    for (int i = 0; i < n && i < max_count; ++i) ms_out[i] = b->solve_ms[i];
    return n;
}

pict_status pict_batch_describe(const pict_batch *b, pict_batch_info *out) {
    PICT_CHECK_ARG(b && out, "NULL argument");
    memset(out, 0, sizeof *out);
    out->count = b->count;
    out->tiles = b->ntiles;
    out->tiles_per_workgroup = b->cfg.tiles_per_workgroup > 0 ? b->cfg.tiles_per_workgroup : 1;
    out->graph = b->last_graph ? 1 : 0;
    if (b->plan) {
        BatchSolveInfo ia, ip;
        batch_solve_describe(b->plan->adv, &ia);
        batch_solve_describe(b->plan->prs, &ip);
        out->advection_grid = ia.grid;
        out->advection_jobs = ia.jobs;
        out->advection_rows_per_workgroup = ia.rows_per_wg;
        out->advection_rows_per_thread = ia.rows_per_thread;
        out->pressure_grid = ip.grid;
        out->pressure_jobs = ip.jobs;
        out->pressure_rows_per_thread = ip.rows_per_thread;
        snprintf(out->advection_kernel, sizeof out->advection_kernel, "%s", ia.kernel);
        snprintf(out->pressure_kernel, sizeof out->pressure_kernel, "%s", ip.kernel);
    }
    return PICT_OK;
}

}  // extern "C"
