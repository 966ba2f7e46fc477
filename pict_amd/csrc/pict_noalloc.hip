// pict_noalloc.hip -- pict_solve_linear without device allocation.
//
// For callers that cannot call hipMalloc (code under HIP graph capture, tight serving loops, harnesses that
// forbid it): the solver workspace and the solve's final states live in the file-scope device array `scratch`
// below instead of the library's workspace pool.  The array is part of the library's code object, so every GPU
// the library runs on holds one copy of kScratchBytes from module load until the process ends.
//
// Every call on a GPU uses that GPU's one copy: calls that overlap on one GPU (two host threads, or two streams
// whose solves run at the same time) race on it and must not be issued.  The regions the solver needs zeroed
// (its control block and tagged partials) are cleared by a fill kernel at the start of each call, so the contents
// the previous call left behind never matter.
#include <hip/hip_runtime.h>

#include <mutex>

#include "pict_host.hpp"
#include "pict_solver.hpp"

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per device
constexpr int kMaxDevices = 64;
constexpr int kMaxRhs = 4;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on each device, looked up once per device with that device current.  In host code
// the array's name is only a stand-in for the symbol, never an address to pass to a kernel or a copy.
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

pict_status scratch_address(void **out) {
    int dev = 0;
    PICT_HIP_TRY(hipGetDevice(&dev));
    PICT_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "device index %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        PICT_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        PICT_CHECK_ARG(p != nullptr, "no device address for the solver scratch array");
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return PICT_OK;
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

extern "C" {

pict_status pict_solve_linear_noalloc(const pict_csr_desc *A, const void *rhs, void *x, int32_t nrhs,
                                      const pict_solve_opts *opts, pict_solve_info *infos_out, void *stream) {
    PICT_CHECK_ARG(A && rhs && x && opts && infos_out, "NULL argument");
    PICT_CHECK_ARG(nrhs >= 1 && nrhs <= kMaxRhs, "nrhs must be 1..%d (got %d)", kMaxRhs, nrhs);
    size_t ws_bytes = 0;
    pict_status st = pict::solve_workspace_bytes(A, nrhs, opts, &ws_bytes);
    if (st) return st;
    // layout: final states | barrier error flag | workspace
    const size_t state_bytes = pict::solve_state_bytes() * (size_t)nrhs;
    const size_t off_err = align256(state_bytes);
    const size_t off_ws = off_err + 256;
    PICT_CHECK_ARG(off_ws + ws_bytes <= kScratchBytes,
                   "pict_solve_linear_noalloc: this solve needs %zu bytes of scratch, more than the %zu bytes it has "
                   "(n = %d, %d right-hand side(s)); use pict_solve_linear",
                   off_ws + ws_bytes, kScratchBytes, A->n, nrhs);
    void *base = nullptr;
    if ((st = scratch_address(&base))) return st;
    char *b = (char *)base;
    hipStream_t s = (hipStream_t)stream;
    st = pict::solve_enqueue(A, rhs, x, nrhs, opts, b, (int *)(b + off_err), nullptr, nullptr, s, b + off_ws,
                             kScratchBytes - off_ws);
    if (st) return st;
    unsigned char host[kMaxRhs * 256];
    int err = 0;
    PICT_CHECK_ARG(state_bytes <= sizeof host, "solver state larger than expected");
    PICT_HIP_TRY(hipMemcpyAsync(host, b, state_bytes, hipMemcpyDeviceToHost, s));
    PICT_HIP_TRY(hipMemcpyAsync(&err, b + off_err, sizeof(int), hipMemcpyDeviceToHost, s));
    PICT_HIP_TRY(hipStreamSynchronize(s));
    if (err) {
        pict::set_error("solver grid barrier exceeded its spin limit");
        return PICT_ERR_DEVICE_SYNC;
    }
    pict::solve_state_to_info(host, nrhs, infos_out);
    return PICT_OK;
}

}  // extern "C"
