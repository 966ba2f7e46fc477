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

}  // namespace pict

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
        return k;
    }
};
