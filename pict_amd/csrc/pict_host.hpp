// pict_host.hpp -- host-side internals shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/pict_hip.h"
#include "pict_device.hpp"

namespace pict {

void set_error(const char *fmt, ...);

// Mean subtraction of a global vector: kMeanBlocks workgroups of kMeanThreads threads form FP64
// partial sums in a fixed order (the single-domain and the batched step use the same shape, so
// their means are bitwise equal).
constexpr int kMeanThreads = 256;
constexpr int kMeanBlocks = 128;

// Stream-ordered workspace memory from a library-owned pool of the current device that keeps
// what is freed, so repeated calls reuse it instead of allocating device memory each time.
hipError_t ws_alloc(void **p, size_t bytes, hipStream_t s);
hipError_t ws_free(void *p, hipStream_t s);

// The ILU(0)-preconditioned BiCGStab (pict_ilu.hip): synchronous, infos written on return.
pict_status ilu_bicgstab_solve(const pict_csr_desc *A, const void *rhs, void *x, int nrhs, const pict_solve_opts *o,
                               pict_solve_info *infos, hipStream_t s);

// Frees the captured step graphs of an atlas (defined in pict_step.hip).
void step_graphs_free(struct pict_atlas *a);

// Fused forward stages of the native step (pict_stages.hip): advection matrix + RHS + velocity result = 0;
// pressure matrix (if `matrix`) + pressure RHS + pressure result = 0.
pict_status stage_advection_setup(const struct pict_atlas *a, double dt, hipStream_t stream);
pict_status stage_pressure_setup(const struct pict_atlas *a, double dt, bool matrix, hipStream_t stream);

#define PICT_HIP_TRY(expr)                                                                  \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess) {                                                             \
            ::pict::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, \
                              __LINE__);                                                    \
            return PICT_ERR_HIP;                                                            \
        }                                                                                   \
    } while (0)

#define PICT_CHECK_ARG(cond, ...)            \
    do {                                     \
        if (!(cond)) {                       \
            ::pict::set_error(__VA_ARGS__);  \
            return PICT_ERR_INVALID;         \
        }                                    \
    } while (0)

inline pict_status check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("kernel launch failed: %s", hipGetErrorString(e));
        return PICT_ERR_HIP;
    }
    return PICT_OK;
}


}  // namespace pict

// dtype x spatial-dimension dispatch of a kernel template <R, D>
#define PICT_DISPATCH(atlas, BODY)                                                   \
    do {                                                                             \
        const int _D = (atlas)->desc.num_dims;                                       \
        if ((atlas)->desc.dtype == PICT_DTYPE_F32) {                                 \
            using R = float;                                                         \
            if (_D == 1) { constexpr int D = 1; BODY; }                              \
            else if (_D == 2) { constexpr int D = 2; BODY; }                         \
            else { constexpr int D = 3; BODY; }                                      \
        } else {                                                                     \
            using R = double;                                                        \
            if (_D == 1) { constexpr int D = 1; BODY; }                              \
            else if (_D == 2) { constexpr int D = 2; BODY; }                         \
            else { constexpr int D = 3; BODY; }                                      \
        }                                                                            \
    } while (0)

// one workgroup per 256-cell tile of a block
#define PICT_LAUNCH_CELLS(atlas, stream, KERNEL, ...)                                           \
    hipLaunchKernelGGL(KERNEL, dim3((atlas)->num_tiles), dim3(kCellsPerTile), 0,                \
                       (hipStream_t)(stream), __VA_ARGS__)

namespace pict {
struct GraphCache;
}

// Device-resident domain: descriptor copy + launch tables (one allocation).
struct pict_atlas {
    pict_domain_desc desc;                // host copy; desc.blocks points to host_blocks
    std::vector<pict_block_desc> host_blocks;
    int num_tiles = 0;
    void *dev = nullptr;                  // [blocks | tile_block | tile_start]
    size_t dev_bytes = 0;
    pict_block_desc *dev_blocks = nullptr;
    int32_t *dev_tile_block = nullptr;
    int32_t *dev_tile_start = nullptr;
    int device = 0;
    // pinned staging ring for stream-ordered block-table uploads
    static constexpr int kRing = 8;
    void *pinned[kRing] = {};
    hipEvent_t ev[kRing] = {};
    int ring_pos = 0;
    // A^T map of the CSR pattern (built on first use by a reverse pass, for the pattern it was built from)
    mutable int32_t *tperm = nullptr;
    mutable const int32_t *tperm_row = nullptr, *tperm_index = nullptr;
    // captured graphs of native steps and reverse passes (pict_step.hip)
    mutable pict::GraphCache *step_graphs = nullptr;

    template <typename R>
    pict::DomainK<R> view() const {
        pict::DomainK<R> k;
        k.D = desc.num_dims;
        k.N = desc.num_cells;
        k.num_blocks = desc.num_blocks;
        k.nnz = desc.nnz;
        k.viscosity = (R)desc.viscosity;
        k.blocks = dev_blocks;
        k.tile_block = dev_tile_block;
        k.tile_start = dev_tile_start;
        k.num_tiles = num_tiles;
        k.csr_row = desc.csr_row;
        k.csr_index = desc.csr_index;
        k.C = (R *)desc.C_value;
        k.P = (R *)desc.P_value;
        k.A = (R *)desc.A;
        k.vrhs = (R *)desc.velocity_rhs;
        k.vres = (R *)desc.velocity_result;
        k.prhs = (R *)desc.pressure_rhs;
        k.prhs_div = (R *)desc.pressure_rhs_div;
        k.pres = (R *)desc.pressure_result;
        k.visc_grad = (R *)desc.viscosity_grad;
        k.C_grad = (R *)desc.C_grad;
        k.P_grad = (R *)desc.P_grad;
        k.A_grad = (R *)desc.A_grad;
        k.vrhs_grad = (R *)desc.velocity_rhs_grad;
        k.vres_grad = (R *)desc.velocity_result_grad;
        k.prhs_grad = (R *)desc.pressure_rhs_grad;
        k.prhs_div_grad = (R *)desc.pressure_rhs_div_grad;
        k.pres_grad = (R *)desc.pressure_result_grad;
        k.channels = desc.scalar_channels;
        k.svisc_static = desc.scalar_viscosity_static;
        k.svisc = (const R *)desc.scalar_viscosity;
        k.svisc_grad = (R *)desc.scalar_viscosity_grad;
        k.srhs = (R *)desc.scalar_rhs;
        k.sres = (R *)desc.scalar_result;
        k.srhs_grad = (R *)desc.scalar_rhs_grad;
        k.sres_grad = (R *)desc.scalar_result_grad;
        return k;
    }
};
