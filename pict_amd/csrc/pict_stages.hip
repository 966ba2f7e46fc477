// pict_stages.hip -- domain atlas and the per-cell PISO stage kernels.
//
// One workgroup = 256 consecutive cells of one block, one thread per cell.
// The workgroup->(block, first cell) tables are built once per domain
// structure and live in device memory next to the block table (the reference
// rebuilds and uploads them on every call, PISO_multiblock_cuda_kernel.cu:59-93).
// The block table is uploaded stream-ordered, so no host synchronisation is
// needed when tensors are swapped between stages.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "pict_cells.hpp"
#include "pict_host.hpp"

namespace pict {

static thread_local std::string g_err;

void set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

// One stream-ordered pool per device, owned by the library, that keeps freed memory (release
// threshold at its maximum): the solver and step workspaces are allocated and freed on every
// call, and a pool that trims at each synchronisation makes each of those a fresh device
// allocation.
static hipMemPool_t device_pool(int dev, hipError_t &err) {
    static std::mutex mu;
    static hipMemPool_t pools[64] = {};
    err = hipSuccess;
    if (dev < 0 || dev >= 64) {
        err = hipErrorInvalidDevice;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(mu);
    if (!pools[dev]) {
        hipMemPoolProps props;
        memset(&props, 0, sizeof props);
        props.allocType = hipMemAllocationTypePinned;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = dev;
        hipMemPool_t pool = nullptr;
        err = hipMemPoolCreate(&pool, &props);
        if (err != hipSuccess) return nullptr;
        uint64_t keep = ~uint64_t(0);
        err = hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        if (err != hipSuccess) {
            (void)hipMemPoolDestroy(pool);
            return nullptr;
        }
        pools[dev] = pool;
    }
    return pools[dev];
}

hipError_t ws_alloc(void **p, size_t bytes, hipStream_t s) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    hipMemPool_t pool = device_pool(dev, e);
    if (!pool) return e;
    return hipMallocFromPoolAsync(p, bytes, pool, s);
}

hipError_t ws_free(void *p, hipStream_t s) { return hipFreeAsync(p, s); }

// ---------------------------------------------------------------------------
// CSR pattern: row pointer from the closed form of
// getCSRMatrixRowEndOffsetFromBlockBoundaries3D (:247-312), columns of
// diag + non-prescribed neighbours sorted ascending (:3663-3772, :3846-3860).
// ---------------------------------------------------------------------------
template <int D>
__device__ int row_end_closed_form(const pict_block_desc *b, int flatPos, int &rs) {
    const I4 p = unflat(flatPos, b);
    rs = 2 * D + 1;
    int re = (flatPos + 1) * rs;
    if (prescribed(0, b)) { re -= flatPos / b->size[0] + 1; if (p.a[0] == 0) --rs; }
    if (prescribed(1, b)) { re -= (flatPos + 1) / b->size[0]; if (p.a[0] == b->size[0] - 1) --rs; }
    if (D > 1) {
        if (prescribed(2, b)) { re -= b->size[0] * p.a[2] + (p.a[1] == 0 ? p.a[0] + 1 : b->size[0]); if (p.a[1] == 0) --rs; }
        if (prescribed(3, b)) { re -= b->size[0] * p.a[2] + (p.a[1] == b->size[1] - 1 ? p.a[0] + 1 : 0); if (p.a[1] == b->size[1] - 1) --rs; }
    }
    if (D > 2) {
        if (prescribed(4, b)) { re -= (p.a[2] == 0 ? flatPos + 1 : b->stride[2]); if (p.a[2] == 0) --rs; }
        if (prescribed(5, b)) { re -= (p.a[2] == b->size[2] - 1 ? flatPos + 1 - p.a[2] * b->stride[2] : 0); if (p.a[2] == b->size[2] - 1) --rs; }
    }
    return re + b->csr_offset;
}

template <typename R, int D>
__global__ __launch_bounds__(kCellsPerTile) void k_csr_pattern(DomainK<R> dom) {
    const pict_block_desc *b;
    int bi, f;
    if (!tile_cell(dom, b, bi, f)) return;
    const I4 pos = unflat(f, b);
    int rs;
    const int re = row_end_closed_form<D>(b, f, rs);
    const int g = b->global_offset + f;
    dom.csr_row[g + 1] = re;
    if (g == 0) dom.csr_row[0] = 0;
    int idx[7];
    row_slots<R, D>(dom, b, bi, pos, idx);
    const int start = re - rs;
    for (int i = 0; i < rs; ++i) {
        const int s = lowest_slot(idx);
        dom.csr_index[start + i] = s < 0 ? -1 : idx[s];
        if (s >= 0) idx[s] = -1;
    }
}

// Forward stage kernels: one workgroup per tile of the domain (bodies in pict_cells.hpp).  NO: the instance with
// the non-orthogonal terms of `flags` (the orthogonal instance carries none of their code).
template <typename R, int D, bool NO>
__global__ __launch_bounds__(kCellsPerTile) void k_adv_matrix(DomainK<R> dom, R dt, int flags) {
    adv_matrix_cell<R, D, NO>(dom, (int)blockIdx.x, dt, flags);
}

template <typename R, int D, bool NO>
__global__ __launch_bounds__(kCellsPerTile) void k_adv_matrix_scalar(DomainK<R> dom, R dt, int flags, int ch) {
    adv_matrix_cell<R, D, NO, true>(dom, (int)blockIdx.x, dt, flags, ch);
}

template <typename R, int D, bool NO>
__global__ __launch_bounds__(kCellsPerTile) void k_scalar_rhs(DomainK<R> dom, R dt, int flags) {
    scalar_rhs_cell<R, D, NO>(dom, (int)blockIdx.x, dt, flags);
}

template <typename R, int D, bool NO>
__global__ __launch_bounds__(kCellsPerTile) void k_adv_rhs(DomainK<R> dom, R dt, int apply_pg, int flags) {
    adv_rhs_cell<R, D, NO>(dom, (int)blockIdx.x, dt, apply_pg, flags);
}

template <typename R, int D, bool NO>
__global__ __launch_bounds__(kCellsPerTile) void k_p_matrix(DomainK<R> dom, int flags) {
    p_matrix_cell<R, D, NO>(dom, (int)blockIdx.x, flags);
}

template <typename R, int D>
__global__ __launch_bounds__(kCellsPerTile) void k_p_rhs_nonortho(DomainK<R> dom, int flags) {
    p_rhs_nonortho_cell<R, D>(dom, (int)blockIdx.x, flags);
}

template <typename R, int D>
__global__ __launch_bounds__(kCellsPerTile) void k_p_rhs(DomainK<R> dom, R dt) {
    p_rhs_cell<R, D>(dom, (int)blockIdx.x, dt);
}

template <typename R, int D>
__global__ __launch_bounds__(kCellsPerTile) void k_divergence(DomainK<R> dom, const R *vg, R *out, R dt, int tsn) {
    divergence_cell<R, D>(dom, (int)blockIdx.x, vg, out, dt, tsn);
}

template <typename R, int D, bool NO>
__global__ __launch_bounds__(kCellsPerTile) void k_p_matrix_face(DomainK<R> dom, int flags) {
    p_matrix_cell<R, D, NO, true>(dom, (int)blockIdx.x, flags);
}

template <typename R, int D>
__global__ __launch_bounds__(kCellsPerTile) void k_p_rhs_nonortho_face(DomainK<R> dom, int flags) {
    p_rhs_nonortho_cell<R, D, true>(dom, (int)blockIdx.x, flags);
}

// the non-orthogonal terms with face transforms read every face's transform of a transformed block
// (getLaplaceCoefficientsSingleFace :1320-1343)
static pict_status check_face_transforms(const pict_atlas *a) {
    for (int i = 0; i < a->desc.num_blocks; ++i) {
        const pict_block_desc &b = a->host_blocks[i];
        PICT_CHECK_ARG(!b.has_transform || (b.has_face_transform && b.face_transform),
                       "useFaceTransform with non-orthogonal terms: block %d has a transform but no face transforms", i);
    }
    return PICT_OK;
}

template <typename R, int D>
__global__ __launch_bounds__(kCellsPerTile) void k_divergence_face(DomainK<R> dom, const R *vg, R *out, R dt,
                                                                   int tsn) {
    divergence_face_cell<R, D>(dom, (int)blockIdx.x, vg, out, dt, tsn);
}

template <typename R, int D>
__global__ __launch_bounds__(kCellsPerTile) void k_correct_velocity(DomainK<R> dom, R dt, int tsn) {
    correct_velocity_cell<R, D>(dom, (int)blockIdx.x, dt, tsn);
}

template <typename R, int D, int MODE>
__global__ __launch_bounds__(kCellsPerTile) void k_correct_velocity_fvm(DomainK<R> dom, R dt, int tsn) {
    correct_velocity_fvm_cell<R, D, MODE>(dom, (int)blockIdx.x, dt, tsn);
}

template <typename R, int D>
__global__ __launch_bounds__(kCellsPerTile) void k_correct_velocity_v4(DomainK<R> dom, R dt) {
    correct_velocity_v4_cell<R, D>(dom, (int)blockIdx.x, dt);
}

template <typename R>
__global__ __launch_bounds__(kCellsPerTile) void k_copy_result(DomainK<R> dom, R *gv, int field, int to_blocks) {
    copy_result_cell<R>(dom, (int)blockIdx.x, gv, field, to_blocks);
}

// Fused stages of the native step (pict_step.hip): stages that read only what earlier launches wrote share one
// launch, with the same values as the separate kernels.
// advection matrix + advection RHS + solve start vector 0
template <typename R, int D>
__global__ __launch_bounds__(kCellsPerTile) void k_advection_setup(DomainK<R> dom, R dt) {
    const int t = (int)blockIdx.x;
    adv_matrix_cell<R, D>(dom, t, dt);
    adv_rhs_cell<R, D>(dom, t, dt, 0);
    zero_result_cell<R>(dom, t, 0);
}
// pressure matrix (when `matrix`: it depends on A only, so once per step) + pressure RHS + solve start vector 0
template <typename R, int D>
__global__ __launch_bounds__(kCellsPerTile) void k_pressure_setup(DomainK<R> dom, R dt, int matrix) {
    const int t = (int)blockIdx.x;
    if (matrix) p_matrix_cell<R, D>(dom, t);
    p_rhs_cell<R, D>(dom, t, dt);
    zero_result_cell<R>(dom, t, 1);
}

}  // namespace pict

using namespace pict;

static pict_status validate_desc(const pict_domain_desc *d) {
    PICT_CHECK_ARG(d != nullptr, "descriptor is NULL");
    PICT_CHECK_ARG(d->num_dims >= 1 && d->num_dims <= 3, "num_dims must be 1..3 (got %d)", d->num_dims);
    PICT_CHECK_ARG(d->num_blocks >= 1 && d->blocks != nullptr, "domain has no blocks");
    PICT_CHECK_ARG(d->dtype == PICT_DTYPE_F32 || d->dtype == PICT_DTYPE_F64, "unknown dtype %d", d->dtype);
    long long cells = 0;
    for (int i = 0; i < d->num_blocks; ++i) {
        const pict_block_desc &b = d->blocks[i];
        PICT_CHECK_ARG(b.global_offset == cells, "block %d: global_offset %d, expected %lld", i, b.global_offset, cells);
        for (int a = 0; a < 3; ++a) {
            PICT_CHECK_ARG(b.size[a] >= 1, "block %d: invalid size", i);
            if (a < d->num_dims) PICT_CHECK_ARG(b.size[a] > 2, "block %d: spatial sizes must be at least 3", i);
        }
        PICT_CHECK_ARG(b.stride[3] == b.size[0] * b.size[1] * b.size[2], "block %d: stride.w mismatch", i);
        PICT_CHECK_ARG(b.velocity && b.pressure, "block %d: velocity/pressure missing", i);
        PICT_CHECK_ARG(!b.has_transform || b.transform, "block %d: transform flag without data", i);
        for (int f = 0; f < 2 * d->num_dims; ++f) {
            const pict_boundary_desc &bd = b.bounds[f];
            if (bd.type == PICT_BOUNDARY_CONNECTED) {
                PICT_CHECK_ARG(bd.connected_block >= 0 && bd.connected_block < d->num_blocks,
                               "block %d face %d: connected block out of range", i, f);
            } else if (bd.type == PICT_BOUNDARY_FIXED) {
                PICT_CHECK_ARG(bd.velocity != nullptr, "block %d face %d: fixed boundary without velocity", i, f);
                PICT_CHECK_ARG(d->scalar_channels == 0 || bd.scalar != nullptr,
                               "block %d face %d: fixed boundary without passive scalar", i, f);
                PICT_CHECK_ARG(!bd.has_transform || bd.transform, "block %d face %d: transform flag without data", i, f);
            } else if (bd.type != PICT_BOUNDARY_PERIODIC) {
                set_error("block %d face %d: unsupported boundary type %d", i, f, bd.type);
                return PICT_ERR_UNSUPPORTED;
            }
        }
        cells += b.stride[3];
    }
    PICT_CHECK_ARG(cells == d->num_cells, "num_cells %d does not match blocks (%lld)", d->num_cells, cells);
    PICT_CHECK_ARG(d->scalar_channels >= 0, "negative passive-scalar channel count");
    return PICT_OK;
}

extern "C" {

const char *pict_last_error(void) { return g_err.c_str(); }
int32_t pict_abi_version(void) { return PICT_ABI_VERSION; }

pict_status pict_atlas_create(const pict_domain_desc *desc, pict_atlas **out, void *stream) {
    PICT_CHECK_ARG(out != nullptr, "out is NULL");
    pict_status st = validate_desc(desc);
    if (st) return st;
    pict_atlas *a = new pict_atlas();
    a->desc = *desc;
    a->host_blocks.assign(desc->blocks, desc->blocks + desc->num_blocks);
    a->desc.blocks = a->host_blocks.data();
    std::vector<int32_t> tb, ts;
    for (int i = 0; i < desc->num_blocks; ++i) {
        const int cells = desc->blocks[i].stride[3];
        for (int s = 0; s < cells; s += kCellsPerTile) {
            tb.push_back(i);
            ts.push_back(s);
        }
    }
    a->num_tiles = (int)tb.size();
    const size_t bytes_blocks = sizeof(pict_block_desc) * desc->num_blocks;
    const size_t bytes_tiles = sizeof(int32_t) * tb.size();
    a->dev_bytes = bytes_blocks + 2 * bytes_tiles;
    hipError_t e = hipGetDevice(&a->device);
    if (e == hipSuccess) e = hipMalloc(&a->dev, a->dev_bytes);
    if (e != hipSuccess) {
        set_error("atlas allocation failed: %s", hipGetErrorString(e));
        delete a;
        return PICT_ERR_HIP;
    }
    a->dev_blocks = (pict_block_desc *)a->dev;
    a->dev_tile_block = (int32_t *)((char *)a->dev + bytes_blocks);
    a->dev_tile_start = a->dev_tile_block + tb.size();
    std::vector<char> staging(a->dev_bytes);
    memcpy(staging.data(), desc->blocks, bytes_blocks);
    memcpy(staging.data() + bytes_blocks, tb.data(), bytes_tiles);
    memcpy(staging.data() + bytes_blocks + bytes_tiles, ts.data(), bytes_tiles);
    e = hipMemcpyAsync(a->dev, staging.data(), a->dev_bytes, hipMemcpyHostToDevice, (hipStream_t)stream);
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) {
        set_error("atlas upload failed: %s", hipGetErrorString(e));
        (void)hipFree(a->dev);
        delete a;
        return PICT_ERR_HIP;
    }
    *out = a;
    return PICT_OK;
}

static void atlas_free(pict_atlas *a) {
    for (int i = 0; i < pict_atlas::kRing; ++i) {
        if (a->ev[i]) (void)hipEventDestroy(a->ev[i]);
        if (a->pinned[i]) (void)hipHostFree(a->pinned[i]);
    }
    if (a->dev) (void)hipFree(a->dev);
    if (a->tperm) (void)hipFree(a->tperm);
    step_graphs_free(a);
    delete a;
}

pict_status pict_atlas_update(pict_atlas *a, const pict_domain_desc *desc, void *stream) {
    PICT_CHECK_ARG(a != nullptr, "atlas is NULL");
    pict_status st = validate_desc(desc);
    if (st) return st;
    PICT_CHECK_ARG(desc->num_blocks == a->desc.num_blocks && desc->num_cells == a->desc.num_cells &&
                       desc->num_dims == a->desc.num_dims && desc->dtype == a->desc.dtype,
                   "domain structure changed; create a new atlas");
    for (int i = 0; i < desc->num_blocks; ++i)
        for (int k = 0; k < 4; ++k)
            PICT_CHECK_ARG(desc->blocks[i].size[k] == a->host_blocks[i].size[k], "block %d size changed", i);
    const size_t bytes = sizeof(pict_block_desc) * desc->num_blocks;
    const bool same_blocks = memcmp(desc->blocks, a->host_blocks.data(), bytes) == 0;
    a->desc = *desc;
    if (!same_blocks) {
        a->host_blocks.assign(desc->blocks, desc->blocks + desc->num_blocks);
        // Stream-ordered upload from a pinned ring slot: kernels queued earlier on
        // `stream` still read the previous table; the slot is reused only after
        // its previous copy has completed.
        const int slot = a->ring_pos;
        a->ring_pos = (a->ring_pos + 1) % pict_atlas::kRing;
        if (!a->pinned[slot]) {
            PICT_HIP_TRY(hipHostMalloc(&a->pinned[slot], bytes, hipHostMallocDefault));
            PICT_HIP_TRY(hipEventCreateWithFlags(&a->ev[slot], hipEventDisableTiming));
        } else {
            PICT_HIP_TRY(hipEventSynchronize(a->ev[slot]));
        }
        memcpy(a->pinned[slot], a->host_blocks.data(), bytes);
        PICT_HIP_TRY(hipMemcpyAsync(a->dev_blocks, a->pinned[slot], bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
        PICT_HIP_TRY(hipEventRecord(a->ev[slot], (hipStream_t)stream));
    }
    a->desc.blocks = a->host_blocks.data();
    return PICT_OK;
}

pict_status pict_atlas_destroy(pict_atlas *a) {
    if (!a) return PICT_OK;
    // make sure no queued upload still reads a pinned slot
    for (int i = 0; i < pict_atlas::kRing; ++i)
        if (a->ev[i]) (void)hipEventSynchronize(a->ev[i]);
    atlas_free(a);
    return PICT_OK;
}

}  // extern "C"

namespace pict {

pict_status stage_advection_setup(const pict_atlas *a, double dt, hipStream_t stream) {
    PICT_CHECK_ARG(a && a->desc.C_value && a->desc.A && a->desc.velocity_rhs && a->desc.velocity_result,
                   "advection tensors missing");
    PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_advection_setup<R, D>), a->view<R>(), (R)dt));
    return check_launch();
}

pict_status stage_pressure_setup(const pict_atlas *a, double dt, bool matrix, hipStream_t stream) {
    PICT_CHECK_ARG(a && a->desc.P_value && a->desc.A && a->desc.pressure_rhs && a->desc.C_value &&
                       a->desc.velocity_result && a->desc.pressure_result,
                   "pressure tensors missing");
    PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_pressure_setup<R, D>), a->view<R>(), (R)dt, matrix ? 1 : 0));
    return check_launch();
}

}  // namespace pict

extern "C" {

pict_status pict_build_csr_pattern(const pict_atlas *a, void *stream) {
    PICT_CHECK_ARG(a && a->desc.csr_row && a->desc.csr_index, "pattern buffers missing");
    PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_csr_pattern<R, D>), a->view<R>()));
    return check_launch();
}

pict_status pict_setup_advection_matrix(const pict_atlas *a, double dt, int32_t flags, void *stream) {
    PICT_CHECK_ARG(a && a->desc.C_value && a->desc.A, "C/A missing");
    if (a->desc.num_dims > 1 && (flags & kNonOrthoMatrix))
        PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_adv_matrix<R, D, true>), a->view<R>(), (R)dt, (int)flags));
    else
        PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_adv_matrix<R, D, false>), a->view<R>(), (R)dt, 0));
    return check_launch();
}

// a domain with passive-scalar channels and, for per-block use, scalar tensors on every block
static pict_status check_scalar(const pict_atlas *a, bool blocks) {
    PICT_CHECK_ARG(a != nullptr, "atlas is NULL");
    PICT_CHECK_ARG(a->desc.scalar_channels > 0, "domain has no passive scalar");
    if (blocks)
        for (int i = 0; i < a->desc.num_blocks; ++i)
            PICT_CHECK_ARG(a->host_blocks[i].scalar != nullptr, "block %d: passive scalar missing", i);
    return PICT_OK;
}

pict_status pict_setup_advection_matrix_scalar(const pict_atlas *a, double dt, int32_t flags, int32_t channel,
                                               void *stream) {
    pict_status st = check_scalar(a, false);
    if (st) return st;
    PICT_CHECK_ARG(a->desc.C_value && a->desc.A, "C/A missing");
    PICT_CHECK_ARG(channel >= 0 && channel < a->desc.scalar_channels, "passive scalar channel %d out of range",
                   channel);
    if (a->desc.num_dims > 1 && (flags & kNonOrthoMatrix))
        PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_adv_matrix_scalar<R, D, true>), a->view<R>(), (R)dt,
                                           (int)flags, (int)channel));
    else
        PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_adv_matrix_scalar<R, D, false>), a->view<R>(), (R)dt, 0,
                                           (int)channel));
    return check_launch();
}

pict_status pict_setup_advection_scalar(const pict_atlas *a, double dt, int32_t flags, void *stream) {
    pict_status st = check_scalar(a, true);
    if (st) return st;
    PICT_CHECK_ARG(a->desc.scalar_rhs, "scalar_rhs missing");
    if (a->desc.num_dims > 1 && (flags & kNonOrthoRhs)) {
        PICT_CHECK_ARG(a->desc.scalar_result, "scalar_result missing (non-orthogonal RHS terms)");
        PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_scalar_rhs<R, D, true>), a->view<R>(), (R)dt, (int)flags));
    } else {
        PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_scalar_rhs<R, D, false>), a->view<R>(), (R)dt, 0));
    }
    return check_launch();
}

pict_status pict_setup_advection_velocity(const pict_atlas *a, double dt, int32_t flags, int32_t apply_pg,
                                          void *stream) {
    PICT_CHECK_ARG(a && a->desc.velocity_rhs, "velocity_rhs missing");
    if (a->desc.num_dims > 1 && (flags & kNonOrthoRhs)) {
        PICT_CHECK_ARG(a->desc.velocity_result && a->desc.A, "velocity_result missing (non-orthogonal RHS terms)");
        PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_adv_rhs<R, D, true>), a->view<R>(), (R)dt, (int)apply_pg,
                                           (int)flags));
    } else {
        PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_adv_rhs<R, D, false>), a->view<R>(), (R)dt, (int)apply_pg, 0));
    }
    return check_launch();
}

pict_status pict_setup_pressure_matrix(const pict_atlas *a, double dt, int32_t flags, int32_t use_face_transform,
                                       void *stream) {
    (void)dt;
    PICT_CHECK_ARG(a && a->desc.P_value && a->desc.A, "P/A missing");
    if (use_face_transform) {
        // useFaceTransform (:4857-4865, :4882-4883): the coefficients from the faces' transforms
        if (a->desc.num_dims > 1 && (flags & kNonOrthoMatrix)) {
            pict_status st = check_face_transforms(a);
            if (st) return st;
            PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_p_matrix_face<R, D, true>), a->view<R>(), (int)flags));
        } else {
            PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_p_matrix_face<R, D, false>), a->view<R>(), 0));
        }
        return check_launch();
    }
    if (a->desc.num_dims > 1 && (flags & kNonOrthoMatrix))
        PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_p_matrix<R, D, true>), a->view<R>(), (int)flags));
    else
        PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_p_matrix<R, D, false>), a->view<R>(), 0));
    return check_launch();
}

pict_status pict_setup_pressure_rhs_div(const pict_atlas *a, double dt, int32_t flags, int32_t use_face_transform,
                                        int32_t tsn, void *stream) {
    PICT_CHECK_ARG(a && a->desc.pressure_rhs && a->desc.pressure_rhs_div, "pressure rhs missing");
    if (use_face_transform) {
        // computeFluxesWithFaceTransforms :1703-1785 reads every face's transform
        for (int i = 0; i < a->desc.num_blocks; ++i)
            PICT_CHECK_ARG(a->host_blocks[i].has_face_transform && a->host_blocks[i].face_transform,
                           "useFaceTransform: block %d has no face transforms", i);
        PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_divergence_face<R, D>), a->view<R>(),
                                           (const R *)a->desc.pressure_rhs, (R *)a->desc.pressure_rhs_div, (R)dt,
                                           (int)tsn));
        pict_status st = check_launch();
        if (st || a->desc.num_dims < 2 || !(flags & kNonOrthoRhs)) return st;
        // k_pressureRHSaddNonOrthoComponents with face coefficients (:5455-5477, :3128-3146)
        PICT_CHECK_ARG(a->desc.pressure_result && a->desc.A, "pressure_result missing (non-orthogonal RHS terms)");
        PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_p_rhs_nonortho_face<R, D>), a->view<R>(), (int)flags));
        return check_launch();
    }
    PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_divergence<R, D>), a->view<R>(),
                                       (const R *)a->desc.pressure_rhs, (R *)a->desc.pressure_rhs_div, (R)dt,
                                       (int)tsn));
    pict_status st = check_launch();
    if (st || a->desc.num_dims < 2 || !(flags & kNonOrthoRhs)) return st;
    // k_pressureRHSaddNonOrthoComponents (:5455-5477, launched by SetupPressureRHSdiv :5659-5704)
    PICT_CHECK_ARG(a->desc.pressure_result && a->desc.A, "pressure_result missing (non-orthogonal RHS terms)");
    PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_p_rhs_nonortho<R, D>), a->view<R>(), (int)flags));
    return check_launch();
}

pict_status pict_setup_pressure_rhs(const pict_atlas *a, double dt, int32_t flags, int32_t use_face_transform,
                                    int32_t tsn, void *stream) {
    PICT_CHECK_ARG(a && a->desc.pressure_rhs && a->desc.C_value && a->desc.velocity_result, "pressure rhs inputs missing");
    pict_status st;
    PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_p_rhs<R, D>), a->view<R>(), (R)dt));
    st = check_launch();
    if (st) return st;
    return pict_setup_pressure_rhs_div(a, dt, flags, use_face_transform, tsn, stream);
}

pict_status pict_setup_pressure_correction(const pict_atlas *a, double dt, int32_t flags,
                                           int32_t use_face_transform, int32_t tsn, void *stream) {
    pict_status st = pict_setup_pressure_matrix(a, dt, flags, use_face_transform, stream);
    if (st) return st;
    return pict_setup_pressure_rhs(a, dt, flags, use_face_transform, tsn, stream);
}

pict_status pict_correct_velocity(const pict_atlas *a, double dt, int32_t version, int32_t tsn, void *stream) {
    PICT_CHECK_ARG(a && a->desc.velocity_result && a->desc.pressure_rhs && a->desc.A, "inputs missing");
    // the version switch of _CorrectVelocity :6163-6203: 4 the orthogonal flux form, 5 and 6 the finite-volume
    // pressure gradient (cell fluxes / face metrics), every other version finite differences
    if (version == 4) {
        PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_correct_velocity_v4<R, D>), a->view<R>(), (R)dt));
        return check_launch();
    }
    if (version == 5 || version == 6) {
        // a transformed block's walls read their boundary transforms (the face-transform slices), and version 6
        // every face's transform: refuse what the kernel would otherwise read as NULL
        for (int i = 0; i < a->desc.num_blocks; ++i) {
            const pict_block_desc &b = a->host_blocks[i];
            if (!b.has_transform) continue;
            PICT_CHECK_ARG(version == 5 || (b.has_face_transform && b.face_transform),
                           "CorrectVelocity version 6: block %d has a transform but no face transforms", i);
            for (int f = 0; f < 2 * a->desc.num_dims; ++f)
                PICT_CHECK_ARG(b.bounds[f].type != PICT_BOUNDARY_FIXED || (b.bounds[f].has_transform &&
                                                                          b.bounds[f].transform),
                               "CorrectVelocity version %d: block %d face %d is a fixed boundary without a transform "
                               "(set the block's face transforms)", version, i, f);
        }
        if (version == 5)
            PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_correct_velocity_fvm<R, D, 0>), a->view<R>(), (R)dt,
                                               (int)tsn));
        else
            PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_correct_velocity_fvm<R, D, 3>), a->view<R>(), (R)dt,
                                               (int)tsn));
        return check_launch();
    }
    PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_correct_velocity<R, D>), a->view<R>(), (R)dt, (int)tsn));
    return check_launch();
}

pict_status pict_compute_velocity_divergence(const pict_atlas *a, void *out, void *stream) {
    PICT_CHECK_ARG(a && out, "output missing");
    PICT_DISPATCH(a, PICT_LAUNCH_CELLS(a, stream, (k_divergence<R, D>), a->view<R>(), (const R *)nullptr, (R *)out,
                                       R(1), 0));
    return check_launch();
}

static pict_status copy_result(const pict_atlas *a, int field, int to_blocks, void *stream) {
    PICT_CHECK_ARG(a != nullptr, "atlas is NULL");
    if (field >= 4) {
        pict_status st = check_scalar(a, false);
        if (st) return st;
    }
    void *gv = field == 0 ? a->desc.velocity_result : field == 1 ? a->desc.pressure_result
             : field == 2 ? a->desc.velocity_result_grad : field == 3 ? a->desc.pressure_result_grad
             : field == 4 ? a->desc.scalar_result : a->desc.scalar_result_grad;
    PICT_CHECK_ARG(gv != nullptr, "global vector missing");
    for (int i = 0; i < a->desc.num_blocks; ++i) {
        const pict_block_desc &b = a->host_blocks[i];
        const void *bp = field == 0 ? b.velocity : field == 1 ? b.pressure : field == 2 ? b.velocity_grad
                       : field == 3 ? b.pressure_grad : field == 4 ? b.scalar : b.scalar_grad;
        PICT_CHECK_ARG(bp != nullptr, "block %d: tensor missing", i);
    }
    if (a->desc.dtype == PICT_DTYPE_F32)
        PICT_LAUNCH_CELLS(a, stream, (k_copy_result<float>), a->view<float>(), (float *)gv, field, to_blocks);
    else
        PICT_LAUNCH_CELLS(a, stream, (k_copy_result<double>), a->view<double>(), (double *)gv, field, to_blocks);
    return check_launch();
}
pict_status pict_copy_velocity_result_to_blocks(const pict_atlas *a, void *s) { return copy_result(a, 0, 1, s); }
pict_status pict_copy_velocity_result_from_blocks(const pict_atlas *a, void *s) { return copy_result(a, 0, 0, s); }
pict_status pict_copy_pressure_result_to_blocks(const pict_atlas *a, void *s) { return copy_result(a, 1, 1, s); }
pict_status pict_copy_pressure_result_from_blocks(const pict_atlas *a, void *s) { return copy_result(a, 1, 0, s); }
pict_status pict_copy_velocity_result_grad_to_blocks(const pict_atlas *a, void *s) { return copy_result(a, 2, 1, s); }
pict_status pict_copy_velocity_result_grad_from_blocks(const pict_atlas *a, void *s) { return copy_result(a, 2, 0, s); }
pict_status pict_copy_pressure_result_grad_from_blocks(const pict_atlas *a, void *s) { return copy_result(a, 3, 0, s); }
pict_status pict_copy_scalar_result_to_blocks(const pict_atlas *a, void *s) { return copy_result(a, 4, 1, s); }
pict_status pict_copy_scalar_result_from_blocks(const pict_atlas *a, void *s) { return copy_result(a, 4, 0, s); }
pict_status pict_copy_scalar_result_grad_to_blocks(const pict_atlas *a, void *s) { return copy_result(a, 5, 1, s); }
pict_status pict_copy_scalar_result_grad_from_blocks(const pict_atlas *a, void *s) { return copy_result(a, 5, 0, s); }

}  // extern "C"
