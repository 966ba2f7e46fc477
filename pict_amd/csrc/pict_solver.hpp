// pict_solver.hpp -- library-internal interface of the Krylov solvers (pict_solve.hip)
// for in-library callers that queue solves without waiting for them.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/pict_hip.h"

namespace pict {

// Bytes of the final state of one right-hand side (opaque to callers).
size_t solve_state_bytes();

// Queues one solve on `stream` (no host synchronisation).  When the stream reaches it,
// state_out holds nrhs final states (solve_state_bytes() each) and *err_out the grid-barrier
// error flag, both in device memory.  ev0/ev1 (may be null) are recorded around the kernel.
// `ws` (ws_bytes >= solve_workspace_bytes) is the solve's workspace; null takes one from the
// library's pool for the duration of the solve.  With a caller workspace, and a transposed
// matrix whose transpose map is given, the queued work allocates nothing and never waits on
// the host, so it can be captured into a graph.
pict_status solve_enqueue(const pict_csr_desc *A, const void *rhs, void *x, int nrhs, const pict_solve_opts *o,
                          void *state_out, int *err_out, hipEvent_t ev0, hipEvent_t ev1, hipStream_t stream,
                          void *ws = nullptr, size_t ws_bytes = 0);

// Workspace bytes of one solve of this system (what solve_enqueue takes from the pool).
pict_status solve_workspace_bytes(const pict_csr_desc *A, int nrhs, const pict_solve_opts *o, size_t *bytes);

// Host copies of final states -> pict_solve_info.
void solve_state_to_info(const void *states, int nrhs, pict_solve_info *out);

}  // namespace pict
