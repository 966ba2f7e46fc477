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

namespace pict {

// ---- batched solves: many independent systems in one cooperative launch ----
struct BatchSolve;

struct BatchSolveConfig {
    int rows_per_thread = 0;       // CG (k_cg_local): rows per thread, 0 = default
    int rows_per_wg = 0;           // streaming BiCGStab: rows per workgroup, 0 = default
    int bicg_rows_per_thread = 0;  // > 0: register-resident BiCGStab (FP32, norm-2 criteria) with these rows per thread
};

struct BatchSolveInfo {
    int grid, rows_per_thread, rows_per_wg, local, jobs;
    double makespan;  // of the schedule, in units of the cost estimates
    char kernel[64];
};

// Plans a batched solve of systems A[s] x[s] = rhs[s] (nrhs right-hand sides each, options o; CG
// when !o->use_bicg, then one right-hand side), allocates its workspaces and uploads its tables.
// `cost` (may be null): the estimated time of each solve, for the schedule.
pict_status batch_solve_create(const pict_csr_desc *A, const void *const *rhs, void *const *x, int count, int nrhs,
                               const pict_solve_opts *o, const BatchSolveConfig &cfg, const double *cost,
                               BatchSolve **out, hipStream_t stream);
void batch_solve_destroy(BatchSolve *b);
// new matrix or vector addresses (same sizes); waits for queued work when the table changes
pict_status batch_solve_update(BatchSolve *b, const pict_csr_desc *A, const void *const *rhs, void *const *x,
                               hipStream_t stream);
// new schedule from measured costs (waits for queued work)
pict_status batch_solve_reschedule(BatchSolve *b, const double *cost, hipStream_t stream);
// Queues the solves (no allocation, no host wait: capturable).  states_out (may be null): count*nrhs
// final states (solve_state_bytes() each); err_out (may be null): nonzero when a barrier timed out.
pict_status batch_solve_enqueue(const BatchSolve *b, void *states_out, int *err_out, hipEvent_t ev0, hipEvent_t ev1,
                                hipStream_t stream);
void batch_solve_describe(const BatchSolve *b, BatchSolveInfo *out);

}  // namespace pict
