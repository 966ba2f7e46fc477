/*
 * pict_hip.h -- C ABI of libpict_hip, the MI355X (gfx950) implementation of
 * the PICT PISO time step.
 *
 * Every entry point takes plain pointers, sizes and POD descriptors; no torch
 * or HIP types appear in a signature (streams are passed as `void*`, a
 * hipStream_t, NULL = default stream).  All entry points return a
 * pict_status; on failure pict_last_error() holds a message for the calling
 * thread.
 *
 * Each entry point names the reference interface it replaces
 * (tum-pbs/PICT, paths relative to the repository root).
 */
#ifndef PICT_HIP_H
#define PICT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* status, enums                                                             */
/* ------------------------------------------------------------------------ */

typedef int32_t pict_status;
#define PICT_OK               0
#define PICT_ERR_INVALID      1  /* invalid argument or descriptor          */
#define PICT_ERR_HIP          2  /* HIP runtime error                       */
#define PICT_ERR_UNSUPPORTED  3  /* option not implemented by this library  */
#define PICT_ERR_DEVICE_SYNC  4  /* device-side barrier spin limit reached  */

#define PICT_DTYPE_F32 0
#define PICT_DTYPE_F64 1

/* Boundary types: values of BoundaryType in extensions/domain_structs_gpu.h:98-114 */
#define PICT_BOUNDARY_CONNECTED 20
#define PICT_BOUNDARY_PERIODIC  21
#define PICT_BOUNDARY_FIXED     30

/* BoundaryConditionType, extensions/domain_structs_gpu.h:140-148 */
#define PICT_BC_DIRICHLET 0
#define PICT_BC_NEUMANN   1

/* non-orthogonal flags, extensions/PISO_multiblock_cuda.h:9-13 */
#define PICT_NON_ORTHO_DIRECT_MATRIX   1
#define PICT_NON_ORTHO_DIRECT_RHS      2
#define PICT_NON_ORTHO_DIAGONAL_MATRIX 4
#define PICT_NON_ORTHO_DIAGONAL_RHS    8
#define PICT_NON_ORTHO_CENTER_MATRIX  16

/* ConvergenceCriterion, extensions/bicgstab_solver.h:11-18 */
#define PICT_CONV_NORM2            0
#define PICT_CONV_NORM2_NORMALIZED 1
#define PICT_CONV_ABS_SUM          2
#define PICT_CONV_ABS_MEAN         3
#define PICT_CONV_ABS_MAX          4

/* ------------------------------------------------------------------------ */
/* domain descriptor                                                         */
/* ------------------------------------------------------------------------ */
/*
 * Layouts (identical to the reference, extensions/domain_structs.cpp):
 *  - block fields NCDHW [1, C, (Z), Y, X]; cell flat index x + y*X + z*X*Y
 *  - global vectors component-major: v[d*N + globalOffset + flat]
 *  - transform [1, (Z), Y, X, 2*D*D+1] = {M row-major, Minv row-major, det}
 *  - face transform [1, D, (Z+1), Y+1, X+1, 2*D*D+1]
 *  - CSR: row int32[N+1], index int32[nnz], value real[nnz]; rows sorted by
 *    ascending global column.
 * `real` is float (dtype F32) or double (dtype F64).  Pointers are device
 * pointers for the HIP library.
 */

typedef struct pict_boundary_desc {
    int32_t type;             /* PICT_BOUNDARY_*                                   */
    int32_t connected_block;  /* CONNECTED: index of the other block               */
    int32_t axes[4];          /* CONNECTED: axes[0] = face of the other block,
                                 axes[1..2] = (axis<<1)|inverted of the other block
                                 for axes (face/2+1)%D and (face/2+2)%D              */
    int32_t velocity_static;  /* FIXED: velocity is [1,D] (else [1,D,(z),y,x])    */
    int32_t velocity_bc;      /* FIXED: PICT_BC_*                                  */
    int32_t has_transform;    /* FIXED: transform slice present                    */
    int32_t reserved;
    int32_t size[4];          /* FIXED: x,y,z (face axis 1), w = D                 */
    int32_t stride[4];        /* FIXED: 1, x, x*y, x*y*z                           */
    void   *velocity;         /* FIXED */
    void   *velocity_grad;    /* FIXED, may be NULL */
    void   *transform;        /* FIXED, [1,(z),y,x,2D^2+1] or NULL */
    /* FIXED faces of a domain with passive-scalar channels (FixedBoundary
       passiveScalar and its types, domain_structs.h:190-260)                   */
    int32_t scalar_static;    /* passive scalar is [1,C] (else [1,C,(z),y,x])     */
    int32_t scalar_bc_mask;   /* bit (c & 31) set: channel c is PICT_BC_NEUMANN
                                 (per-channel types for at most 32 channels)     */
    void   *scalar;           /* NULL when the domain has no passive scalar       */
    void   *scalar_grad;      /* may be NULL                                      */
} pict_boundary_desc;

typedef struct pict_block_desc {
    int32_t global_offset;    /* first global cell index of the block              */
    int32_t csr_offset;       /* first CSR entry of the block                      */
    int32_t size[4];          /* x,y,z, w = D                                      */
    int32_t stride[4];        /* 1, x, x*y, x*y*z                                  */
    int32_t viscosity_static; /* viscosity tensor is [1] (else [1,1,(z),y,x])      */
    int32_t velocity_source_static; /* velocity source is [1,D] (else NCDHW)       */
    int32_t has_transform;
    int32_t has_face_transform;
    void *velocity;           /* [1,D,(z),y,x]                                     */
    void *pressure;           /* [1,1,(z),y,x]                                     */
    void *viscosity;          /* NULL -> domain viscosity                          */
    void *velocity_source;    /* NULL -> none                                      */
    void *transform;          /* NULL if !has_transform                            */
    void *face_transform;     /* NULL if !has_face_transform                       */
    void *velocity_grad;
    void *pressure_grad;
    void *viscosity_grad;
    void *velocity_source_grad;
    pict_boundary_desc bounds[6]; /* -x,+x,-y,+y,-z,+z                             */
    void *scalar;             /* passive scalar [1,C,(z),y,x] or NULL              */
    void *scalar_grad;        /* may be NULL                                        */
} pict_block_desc;

typedef struct pict_domain_desc {
    int32_t num_dims;
    int32_t num_blocks;
    int32_t num_cells;        /* N                                                 */
    int32_t nnz;              /* CSR entries of C and P                            */
    int32_t dtype;            /* PICT_DTYPE_*                                      */
    int32_t reserved;
    double  viscosity;        /* domain viscosity (a CPU scalar in the reference)  */
    const pict_block_desc *blocks; /* host array of num_blocks descriptors         */
    int32_t *csr_row;         /* pattern shared by C, P and their gradients        */
    int32_t *csr_index;
    void *C_value;
    void *P_value;
    void *A;                  /* [N] diagonal of C                                 */
    void *velocity_rhs;       /* [D*N] */
    void *velocity_result;    /* [D*N] */
    void *pressure_rhs;       /* [D*N] */
    void *pressure_rhs_div;   /* [N]   */
    void *pressure_result;    /* [N]   */
    void *viscosity_grad;     /* [1] or NULL */
    void *C_grad;             /* [nnz] or NULL */
    void *P_grad;
    void *A_grad;
    void *velocity_rhs_grad;
    void *velocity_result_grad;
    void *pressure_rhs_grad;
    void *pressure_rhs_div_grad;
    void *pressure_result_grad;
    int32_t scalar_channels;          /* passive-scalar channels C (0: none)        */
    int32_t scalar_viscosity_static;  /* scalar_viscosity is [1] (else [C])         */
    void *scalar_viscosity;   /* device [1] or [C]; NULL: the domain viscosity      */
    void *scalar_viscosity_grad;      /* same shape, may be NULL                    */
    void *scalar_rhs;         /* [C*N] */
    void *scalar_result;      /* [C*N] */
    void *scalar_rhs_grad;
    void *scalar_result_grad;
} pict_domain_desc;

/* Device-resident copy of a domain descriptor plus its launch tables.
   Replaces the reference's DomainAtlasSet / CopyDomainToGPU
   (extensions/domain_structs_gpu.h:376-384, PISO_multiblock_cuda_kernel.cu:54-57)
   and its per-call CopyBlockIndices (PISO_multiblock_cuda_kernel.cu:86-93). */
typedef struct pict_atlas pict_atlas;

/* ------------------------------------------------------------------------ */
/* library / atlas                                                           */
/* ------------------------------------------------------------------------ */

const char *pict_last_error(void);
/* Layout version of the descriptor structs above (2: passive-scalar fields). */
#define PICT_ABI_VERSION 2
int32_t     pict_abi_version(void);

/* Domain::PrepareSolve -> SetupDomainGPU (extensions/domain_structs.cpp:2554-2676, :2992-3029) */
pict_status pict_atlas_create(const pict_domain_desc *desc, pict_atlas **out, void *stream);
/* Domain::UpdateDomainData -> UpdateDomainGPU (extensions/domain_structs.cpp:3031-3266).
   The block table is copied on `stream` (stream ordered: kernels queued
   earlier on the same stream still see the previous table).  The structure
   (block count and sizes) must be unchanged since pict_atlas_create. */
pict_status pict_atlas_update(pict_atlas *atlas, const pict_domain_desc *desc, void *stream);
pict_status pict_atlas_destroy(pict_atlas *atlas);

/* Writes the CSR pattern (row, index) of C/P into desc->csr_row/csr_index.
   The reference writes it inside PISO_build_matrix / PISO_build_pressure_matrix
   (extensions/PISO_multiblock_cuda_kernel.cu:3634-3637, :3846-3860); the pattern
   is step invariant, so it is written once per structure here. */
pict_status pict_build_csr_pattern(const pict_atlas *atlas, void *stream);

/* ------------------------------------------------------------------------ */
/* forward stages                                                            */
/* ------------------------------------------------------------------------ */

/* SetupAdvectionMatrix (extensions/PISOtorch.cpp:503-505; kernel
   PISO_multiblock_cuda_kernel.cu:3601-3865): writes C values and A. */
pict_status pict_setup_advection_matrix(const pict_atlas *atlas, double time_step,
                                        int32_t non_ortho_flags, void *stream);
/* SetupAdvectionVelocity (PISOtorch.cpp:509-511; kernel :4281-4385): velocity_rhs. */
pict_status pict_setup_advection_velocity(const pict_atlas *atlas, double time_step,
                                          int32_t non_ortho_flags,
                                          int32_t apply_pressure_gradient, void *stream);
/* SetupPressureMatrix (PISOtorch.cpp:516-518; kernel :4797-4963): P values.  use_face_transform: the coefficients
   from the faces' transforms (:4857-4865, non-orthogonal :1962-1972; a transformed block needs face transforms for
   the non-orthogonal terms, else PICT_ERR_INVALID). */
pict_status pict_setup_pressure_matrix(const pict_atlas *atlas, double time_step,
                                       int32_t non_ortho_flags, int32_t use_face_transform,
                                       void *stream);
/* SetupPressureRHS (PISOtorch.cpp:519-521; kernels :5121-5240 and :5374-5419):
   pressure_rhs and pressure_rhs_div. */
pict_status pict_setup_pressure_rhs(const pict_atlas *atlas, double time_step,
                                    int32_t non_ortho_flags, int32_t use_face_transform,
                                    int32_t time_step_norm, void *stream);
/* SetupPressureRHSdiv (PISOtorch.cpp:522-524; kernel :5374-5419): pressure_rhs_div.  use_face_transform: fluxes
   from the faces' transforms (computeFluxesWithFaceTransforms :1703-1785; every block needs face transforms, else
   PICT_ERR_INVALID) and face coefficients in the non-orthogonal terms (:3128-3146). */
pict_status pict_setup_pressure_rhs_div(const pict_atlas *atlas, double time_step,
                                        int32_t non_ortho_flags, int32_t use_face_transform,
                                        int32_t time_step_norm, void *stream);
/* SetupPressureCorrection (PISOtorch.cpp:512-514; host :5498-5563):
   pressure matrix + pressure RHS + divergence. */
pict_status pict_setup_pressure_correction(const pict_atlas *atlas, double time_step,
                                           int32_t non_ortho_flags, int32_t use_face_transform,
                                           int32_t time_step_norm, void *stream);
/* CorrectVelocity (PISOtorch.cpp:525-527; version switch :6163-6203): version 4 the orthogonal flux form
   (kernel :6052-6160), 5 and 6 the finite-volume pressure gradient with cell fluxes / face metrics (kernel
   :6013-6049, gradient :991-1124; a transformed block needs boundary transforms on its fixed faces, and face
   transforms for 6, else PICT_ERR_INVALID), every other version finite differences (kernel :5947-5980). */
pict_status pict_correct_velocity(const pict_atlas *atlas, double time_step, int32_t version,
                                  int32_t time_step_norm, void *stream);
/* ComputeVelocityDivergence (PISOtorch.cpp:528; kernel :6333-6354): out [N]. */
pict_status pict_compute_velocity_divergence(const pict_atlas *atlas, void *out, void *stream);
/* ComputeSpatialVelocityGradients (PISOtorch.cpp:530; kernel :6445-6467, getBlockDataGradient :2982-3025):
   out [D][D][N], out[(c*D + i)*N + g] = d u_c / d x_i at global cell g (component-major per velocity component). */
pict_status pict_compute_spatial_velocity_gradients(const pict_atlas *atlas, void *out, void *stream);
/* SGSviscosityIncompressibleSmagorinsky (PISOtorch.cpp:540-541; kernel :6898-6952): out [N], the Smagorinsky
   viscosity coefficient * Delta^2 * |S| of every cell. */
pict_status pict_sgs_viscosity_smagorinsky(const pict_atlas *atlas, double coefficient, void *out, void *stream);

/* Copy{Velocity,Pressure}Result{To,From}Blocks (PISOtorch.cpp:533-538; :6541-6733). */
pict_status pict_copy_velocity_result_to_blocks(const pict_atlas *atlas, void *stream);
pict_status pict_copy_velocity_result_from_blocks(const pict_atlas *atlas, void *stream);
pict_status pict_copy_pressure_result_to_blocks(const pict_atlas *atlas, void *stream);
pict_status pict_copy_pressure_result_from_blocks(const pict_atlas *atlas, void *stream);

/* Passive scalars.  SetupAdvectionMatrix(forPassiveScalar=True, passiveScalarChannel)
   (PISOtorch.cpp:503-505; kernel PISO_build_matrix :3601-3865, host :4510-4560): the
   advection matrix of channel `channel` (scalar viscosity, the channel's boundary types)
   into C values and A. */
pict_status pict_setup_advection_matrix_scalar(const pict_atlas *atlas, double time_step,
                                               int32_t non_ortho_flags, int32_t channel, void *stream);
/* SetupAdvectionScalar (PISOtorch.cpp:506-508; kernel kPISO_build_scalar_advection_RHS
   :4080-4190, host :4590-4623): scalar_rhs of every channel. */
pict_status pict_setup_advection_scalar(const pict_atlas *atlas, double time_step,
                                        int32_t non_ortho_flags, void *stream);
/* CopyScalarResult{To,From}Blocks (PISOtorch.cpp:533-534; :6543-6610). */
pict_status pict_copy_scalar_result_to_blocks(const pict_atlas *atlas, void *stream);
pict_status pict_copy_scalar_result_from_blocks(const pict_atlas *atlas, void *stream);

/* ------------------------------------------------------------------------ */
/* backward stages (vector-Jacobian products)                                */
/* ------------------------------------------------------------------------ */
/* Inputs are the *_grad tensors of the stage outputs; outputs are the
   *_grad tensors of its inputs, which the caller zero-initialises (the
   reference's Create*Grad).  Tensors a stage accumulates into are marked +=,
   tensors it overwrites are marked =.  With non-orthogonal flags the stages also differentiate the explicit
   terms: SetupAdvectionVelocityGrad into velocity_result_grad (+=), SetupPressureRHSdivGrad (and the calls that
   include it) into pressure_result_grad and A_grad (+=). */

/* SetupAdvectionMatrixGrad (PISOtorch.cpp:597-599; kernel PISO_build_matrix_GRAD
   :3869-4075): C_grad, A_grad -> block velocity_grad (+=), viscosity grads (+=). */
pict_status pict_setup_advection_matrix_grad(const pict_atlas *atlas, double time_step,
                                             int32_t non_ortho_flags, void *stream);
/* SetupAdvectionVelocityGrad (PISOtorch.cpp:603-605; kernel :4388-4476):
   velocity_rhs_grad -> block velocity_grad (=), velocity_source_grad (+=),
   fixed-boundary velocity_grad (+=), viscosity grads (+=). */
pict_status pict_setup_advection_velocity_grad(const pict_atlas *atlas, double time_step,
                                               int32_t non_ortho_flags,
                                               int32_t apply_pressure_gradient, void *stream);
/* SetupPressureMatrixGrad (PISOtorch.cpp:609-610; kernel :4967-5115): P_grad -> A_grad (+=). */
pict_status pict_setup_pressure_matrix_grad(const pict_atlas *atlas, double time_step,
                                            int32_t non_ortho_flags, int32_t use_face_transform,
                                            void *stream);
/* SetupPressureRHSdivGrad (PISOtorch.cpp:613-614; kernel :5423-5452):
   pressure_rhs_div_grad -> pressure_rhs_grad (+=), fixed-boundary velocity_grad (+=). */
pict_status pict_setup_pressure_rhs_div_grad(const pict_atlas *atlas, double time_step,
                                             int32_t non_ortho_flags, int32_t use_face_transform,
                                             int32_t time_step_norm, void *stream);
/* SetupPressureRHSGrad (PISOtorch.cpp:611-612; host :5808-5844, kernel :5243-5369):
   the divergence VJP above, then pressure_rhs_grad -> block velocity_grad (=),
   velocity_result_grad (+=), C_grad (= on every row), A_grad (+=), velocity
   source, fixed-boundary velocity and viscosity grads (+=). */
pict_status pict_setup_pressure_rhs_grad(const pict_atlas *atlas, double time_step,
                                         int32_t non_ortho_flags, int32_t use_face_transform,
                                         int32_t time_step_norm, void *stream);
/* SetupPressureCorrectionGrad (PISOtorch.cpp:607-608; host :5707-5752):
   pressure RHS grad, then pressure matrix grad. */
pict_status pict_setup_pressure_correction_grad(const pict_atlas *atlas, double time_step,
                                                int32_t non_ortho_flags, int32_t use_face_transform,
                                                int32_t time_step_norm, void *stream);
/* CorrectVelocityGrad (PISOtorch.cpp:615; kernel :6250-6293): velocity_result_grad ->
   pressure_rhs_grad (=), block pressure_grad (+=), A_grad (=). */
pict_status pict_correct_velocity_grad(const pict_atlas *atlas, double time_step,
                                       int32_t time_step_norm, void *stream);
/* Copy{Velocity,Pressure}ResultGrad{To,From}Blocks (PISOtorch.cpp:619-621; :6819-6915). */
pict_status pict_copy_velocity_result_grad_to_blocks(const pict_atlas *atlas, void *stream);
pict_status pict_copy_velocity_result_grad_from_blocks(const pict_atlas *atlas, void *stream);
pict_status pict_copy_pressure_result_grad_from_blocks(const pict_atlas *atlas, void *stream);
/* SetupAdvectionMatrixGrad(forPassiveScalar=True) (PISOtorch.cpp:568-570; kernel
   PISO_build_matrix_GRAD :3869-4075): C_grad, A_grad -> block velocity_grad (+=),
   fixed-boundary velocity_grad (+=), scalar viscosity grad (+=; the domain
   viscosity grad without a scalar viscosity). */
pict_status pict_setup_advection_matrix_scalar_grad(const pict_atlas *atlas, double time_step,
                                                    int32_t non_ortho_flags, int32_t channel, void *stream);
/* SetupAdvectionScalarGrad (PISOtorch.cpp:571-573; kernel :4193-4278, host :4626-4659): scalar_rhs_grad ->
   block scalar_grad (=), fixed-boundary scalar_grad and velocity_grad (+=), scalar
   viscosity grad (+=), scalar_result_grad (+=, explicit non-orthogonal terms). */
pict_status pict_setup_advection_scalar_grad(const pict_atlas *atlas, double time_step,
                                             int32_t non_ortho_flags, void *stream);
/* CopyScalarResultGrad{From,To}Blocks (PISOtorch.cpp:588-589; :6737-6800). */
pict_status pict_copy_scalar_result_grad_from_blocks(const pict_atlas *atlas, void *stream);
pict_status pict_copy_scalar_result_grad_to_blocks(const pict_atlas *atlas, void *stream);

/* ------------------------------------------------------------------------ */
/* linear solvers                                                            */
/* ------------------------------------------------------------------------ */

typedef struct pict_csr_desc {
    int32_t n;                /* rows (= columns)                                  */
    int32_t nnz;
    int32_t dtype;            /* PICT_DTYPE_*                                      */
    int32_t reserved;
    const int32_t *row;       /* [n+1] */
    const int32_t *index;     /* [nnz] */
    const void    *value;     /* [nnz] */
    const int32_t *transpose_perm; /* optional [nnz]: position of (j,i) for entry (i,j), from
                                      pict_csr_transpose_permutation; NULL = built per transposed solve */
} pict_csr_desc;

typedef struct pict_solve_opts {
    int32_t max_iterations;
    int32_t convergence_criterion;   /* PICT_CONV_* */
    int32_t use_bicg;
    int32_t matrix_rank_deficient;   /* must be 0 */
    int32_t residual_reset_steps;    /* CG: recompute r = b - A x every k its; 0 = off */
    int32_t transpose_a;             /* solve A^T x = b */
    int32_t print_residual;
    int32_t return_best_result;      /* CG only */
    int32_t bicg_with_preconditioner;/* BiCGStab with the ILU(0) preconditioner M = L U
                                        (bicgstab_solver_kernel.cu:139-365): pict_solve_linear and
                                        pict_solve_refined only (host analysis, synchronous); the
                                        enqueue-only entry points return UNSUPPORTED */
    int32_t zero_rhs_shortcut;       /* if every right-hand side is exactly 0: x = 0,
                                        no iterations (PISOtorch_diff.py:251, :292-293) */
    double  tolerance;
} pict_solve_opts;

typedef struct pict_solve_info {     /* LinearSolverResultInfo, bicgstab_solver.h:20-25 */
    double  final_residual;
    int32_t used_iterations;
    int32_t converged;
    int32_t is_finite_residual;
    int32_t reserved;
} pict_solve_info;

/* SolveLinear (PISOtorch.cpp:559-563; PISO_multiblock_cuda_kernel.cu:7021-7103,
   cg_solver_kernel.cu:114-456, bicgstab_solver_kernel.cu:47-395).
   rhs and x hold nrhs consecutive vectors of length n; x is the initial guess
   and receives the solution.  infos_out receives nrhs entries (host memory).
   Every right-hand side keeps its own iteration state and stopping test, as
   when the reference solves them one after the other. */
pict_status pict_solve_linear(const pict_csr_desc *A, const void *rhs, void *x, int32_t nrhs,
                              const pict_solve_opts *opts, pict_solve_info *infos_out,
                              void *stream);

/* Name of the device kernel pict_solve_linear would launch for this system
   (as a profiler shows it, e.g. "pict::k_cg_local<float, 4, 5, false>"),
   written NUL-terminated into buf[len].  No counterpart in the reference:
   it labels measurements of the solver kernel. */
/* Mixed-precision solve of A x = b to a relative FP64 residual: x (FP64, initial guess in,
   solution out; A and b in the matrix dtype).  Each sweep forms r = b - A x in FP64, solves
   A d = r / rms(r) in the matrix dtype with `inner` (the CG/BiCGStab kernels of
   pict_solve_linear) and adds rms(r) * d to x, until ||r||_2 / ||b||_2 < refine_tol, the
   residual stops halving, or max_sweeps.  null_space: 1 = the constant vector spans A's null
   space (pure Neumann / periodic pressure systems): the constant part of b and of every
   residual is projected out, since no x can reduce it; 0 = none; -1 = detect it from the row
   sums.  info_out: final relative residual, inner iterations in total, converged, finite;
   `reserved` = sweeps.  No counterpart in the reference, whose pressure solve is one solve in
   the matrix dtype: this is the FP64 refinement of the FP32 pressure solve. */
pict_status pict_solve_refined(const pict_csr_desc *A, const void *rhs, double *x, const pict_solve_opts *inner,
                               double refine_tol, int32_t max_sweeps, int32_t null_space, pict_solve_info *info_out,
                               void *stream);

pict_status pict_solver_kernel_name(const pict_csr_desc *A, int32_t nrhs, const pict_solve_opts *opts, char *buf,
                                    int32_t len);

/* pict_solve_linear without device allocation.  No counterpart in the reference: it is for callers that cannot
   allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
   Same arguments, results and errors as pict_solve_linear; the solver workspace and the final states live in a
   64 MiB device array of the library (one per GPU, held until the process ends) instead of the workspace pool.
   A solve whose workspace exceeds it (about 84 bytes per row for one FP32 right-hand side) returns
   PICT_ERR_INVALID.  Calls that overlap on one GPU share that array and race: issue them one at a time. */
pict_status pict_solve_linear_noalloc(const pict_csr_desc *A, const void *rhs, void *x, int32_t nrhs,
                                      const pict_solve_opts *opts, pict_solve_info *infos_out, void *stream);

/* Transpose map of a structurally symmetric pattern (no counterpart in the reference, which
   transposes inside cuSPARSE): perm_out[e] = position of entry (j,i) for entry e = (i,j).
   Returns PICT_ERR_UNSUPPORTED if some (j,i) is missing.  Synchronises `stream`. */
pict_status pict_csr_transpose_permutation(const pict_csr_desc *pattern, int32_t *perm_out, void *stream);

/* ------------------------------------------------------------------------ */
/* whole step                                                                */
/* ------------------------------------------------------------------------ */

typedef struct pict_step_opts {
    double  time_step;
    int32_t corrector_steps;
    int32_t velocity_corrector;   /* CorrectVelocity version: 0 or 1                    */
    int32_t time_step_norm;
    int32_t normalize_pressure;   /* subtract the mean of every pressure result         */
    pict_solve_opts advection;    /* the velocity solve (num_dims right-hand sides)     */
    pict_solve_opts pressure;     /* each pressure solve                                */
} pict_step_opts;

/* Intermediates of one step for its reverse pass (device buffers of the domain dtype, K = corrector_steps). */
typedef struct pict_step_tape {
    void *C_value;          /* [nnz]     advection matrix values                        */
    void *A;                /* [N]       its diagonal                                   */
    void *P_value;          /* [nnz]     pressure matrix values                         */
    void *velocity_star;    /* [D*N]     advection solve result                         */
    void *pressure_rhs;     /* [K][D*N]  pressure RHS of every corrector                */
    void *pressure_raw;     /* [K][N]    pressure solve results before the mean subtraction */
    void *velocity_result;  /* [K][D*N]  velocity result after every correction         */
} pict_step_tape;

/* One PISO step, orthogonal branch of the reference driver without hooks or passive
   scalars (PISOtorch_simulation.py:1054-1327): the stages and solves of
   PISOtorch_simulation._PISO_split_step queued back to back on `stream`, with one host
   wait at the end.  Block velocities are read and overwritten; block pressures and the
   domain's C, A, P, RHS and result vectors are written.  infos_out receives num_dims
   advection entries followed by one entry per corrector.  No counterpart in the
   reference, whose driver calls each stage from Python.  tape (may be NULL) receives the
   intermediates pict_piso_step_vjp needs. */
pict_status pict_piso_step(const pict_atlas *atlas, const pict_step_opts *opts, const pict_step_tape *tape,
                           pict_solve_info *infos_out, void *stream);
/* Reverse pass of one pict_piso_step recorded on `tape` (no counterpart in the reference,
   whose backward is PISOtorch_diff's autograd graph, PISOtorch_diff.py:295-1234; this is
   the same sequence of stage VJPs, adjoint solves on A^T and -lambda x^T products).
   velocity_grad [D*N, global layout]: in: gradient w.r.t. the step's output block velocities;
   out: gradient w.r.t. its input block velocities.  Velocity-source, viscosity and fixed-
   boundary velocity gradients are accumulated (+=) into the descriptor's gradient tensors.
   Every other gradient tensor of the descriptor (C, P, A, the RHS and result vectors, block
   velocity and pressure) must be present; they, the domain's C, A, P, RHS and result vectors and
   the block pressures are overwritten.  infos_out: num_dims + corrector_steps adjoint solves. */
pict_status pict_piso_step_vjp(const pict_atlas *atlas, const pict_step_opts *opts, const pict_step_tape *tape,
                               void *velocity_grad, pict_solve_info *infos_out, void *stream);
/* pict_piso_step and pict_piso_step_vjp without the host wait (no counterpart in the reference, whose driver
   reads every solve's result on the host before going on): the work is queued on `stream` and *ticket names the
   pending step; pict_step_wait(ticket, infos_out) waits for it and returns the same infos and errors as the
   waiting call would have.  At most 16 steps may be pending per process; a ticket may be collected from any
   thread.  Solver timing (pict_set_step_timing) must be off. */
pict_status pict_piso_step_async(const pict_atlas *atlas, const pict_step_opts *opts, const pict_step_tape *tape,
                                 int32_t *ticket, void *stream);
pict_status pict_piso_step_vjp_async(const pict_atlas *atlas, const pict_step_opts *opts,
                                     const pict_step_tape *tape, void *velocity_grad, int32_t *ticket,
                                     void *stream);
pict_status pict_step_wait(int32_t ticket, pict_solve_info *infos_out);
/* v -= mean(v) in place, the mean accumulated in FP64 in a fixed order: the pressure
   normalisation of the driver (PISOtorch_simulation.py:1243-1244, a torch expression there). */
pict_status pict_subtract_mean(void *v, int32_t n, int32_t dtype, void *stream);
/* Diagnostics: HIP-event kernel times of the solves of the last pict_piso_step. */
pict_status pict_set_step_timing(int32_t enable);
int32_t     pict_last_step_solve_ms(double *ms_out, int32_t max_count);

/* ------------------------------------------------------------------------ */
/* batches of independent simulations (no counterpart in the reference,     */
/* which steps one Domain per PISOtorch_simulation.Simulation)              */
/* ------------------------------------------------------------------------ */

/* A batch steps `count` independent domains (same dtype and dimension; any sizes,
   viscosities and boundaries) together: each stage is one launch over the tiles of every
   domain, each solve one cooperative launch in which every resident workgroup runs a list
   of (system, workgroup) jobs, and the step is captured into a HIP graph and replayed.
   Every domain gets the result of pict_piso_step on it alone, up to the reduction order
   of its Krylov dot products (the workgroup count of a system may differ). */
typedef struct pict_batch pict_batch;

/* Launch shapes (0: default).  The autotuner (pict_batch_tune) picks them by timing. */
typedef struct pict_batch_config {
    int32_t tiles_per_workgroup;        /* 256-cell tiles per workgroup of the stage kernels: 1, 2 or 4 */
    int32_t pressure_rows_per_thread;   /* CG: matrix rows each thread keeps in registers: 1, 2, 4, 8   */
    int32_t advection_rows_per_workgroup; /* BiCGStab: rows per workgroup (>= 64)                     */
    int32_t adaptive_schedule;          /* 1: after the first step, schedule the solves by the
                                           iteration counts it took; 0: by system size only         */
    int32_t advection_rows_per_thread;  /* BiCGStab kept in registers, this many rows per thread: 1, 2,
                                           4 (FP32, norm-2 criteria, up to 2 right-hand sides); -1: the
                                           streaming kernel; 0: default (2 where eligible)           */
    int32_t reserved;
} pict_batch_config;

typedef struct pict_batch_info {
    int32_t count, tiles, tiles_per_workgroup;
    int32_t advection_grid, advection_jobs, advection_rows_per_workgroup, advection_rows_per_thread;
    int32_t pressure_grid, pressure_jobs, pressure_rows_per_thread;
    int32_t graph;                      /* the last step was a graph replay                          */
    char advection_kernel[64];
    char pressure_kernel[64];
} pict_batch_info;

/* The atlases stay owned by the caller and must outlive the batch. */
pict_status pict_batch_create(const pict_atlas *const *atlases, int32_t count, pict_batch **out);
pict_status pict_batch_destroy(pict_batch *batch);
pict_status pict_batch_set_config(pict_batch *batch, const pict_batch_config *config);
pict_status pict_batch_get_config(const pict_batch *batch, pict_batch_config *config);
/* One PISO step of every domain (pict_piso_step's sequence).  time_steps: count entries (NULL:
   opts->time_step for all).  infos_out: count * (num_dims + corrector_steps) entries, domain by
   domain in pict_piso_step's order. */
pict_status pict_batch_step(pict_batch *batch, const pict_step_opts *opts, const double *time_steps,
                            pict_solve_info *infos_out, void *stream);
/* Autotuner: runs the stages and solves of a step on the batch's current state, timing every
   candidate launch shape `repeats` times (the block velocities and pressures are not changed;
   the domains' matrices, RHS and result vectors are), and keeps the fastest; *chosen (may be
   NULL) receives it. */
pict_status pict_batch_tune(pict_batch *batch, const pict_step_opts *opts, const double *time_steps,
                            int32_t repeats, pict_batch_config *chosen, void *stream);
pict_status pict_batch_describe(const pict_batch *batch, pict_batch_info *out);
/* Diagnostics: when enabled, steps run without the graph and record HIP events around every solve
   launch; the kernel times of the last step's solves (advection, then one per corrector, in
   milliseconds) are then available. */
pict_status pict_batch_set_timing(pict_batch *batch, int32_t enable);
int32_t     pict_batch_last_solve_ms(const pict_batch *batch, double *ms_out, int32_t max_count);

/* Diagnostics (no counterpart in the reference): when enabled, the solver
   records HIP events around its kernel launch on the launch stream; the
   elapsed time of the last solve's kernel is then available (milliseconds;
   after pict_solve_refined, the sum over its correction solves). */
pict_status pict_set_kernel_timing(int32_t enable);
double      pict_last_solve_kernel_ms(void);

/* SparseOuterProduct (PISOtorch.cpp:564; cg_solver_kernel.cu:869-908):
   out_value[k] = a[i] * b[j] for every pattern entry k = (i, j). */
pict_status pict_sparse_outer_product(const void *a, const void *b, const pict_csr_desc *pattern,
                                      void *out_value, void *stream);

/* ------------------------------------------------------------------------ */
/* per-cell vector / matrix operations (channel dimension of NCDHW tensors)  */
/* ------------------------------------------------------------------------ */
/* A tensor [1, C, (Z), Y, X] of `cells` = X*Y*Z cells holds per cell a vector (C = num_dims)
   or a row-major num_dims x num_dims matrix (C = num_dims^2); channel c of cell i at c*cells + i. */
#define PICT_MATMUL_TRANSPOSE_A   1
#define PICT_MATMUL_INVERT_A      2
#define PICT_MATMUL_TRANSPOSE_B   4
#define PICT_MATMUL_INVERT_B      8
#define PICT_MATMUL_TRANSPOSE_OUT 16
#define PICT_MATMUL_INVERT_OUT    32

/* matmul (PISOtorch.cpp:604-611; matrix_vector_ops.cu:113-200): out = A B per cell, out a
   matrix (both matrices), a vector (one matrix) or a scalar (two vectors).  Flags apply to
   matrices only: an operand is loaded (transposed), then inverted; the output is inverted,
   then stored (transposed).  Matrix x vector and vector x matrix ignore the output flags. */
pict_status pict_matmul(const void *a, int32_t a_is_matrix, const void *b, int32_t b_is_matrix, void *out,
                        int32_t num_dims, int64_t cells, int32_t dtype, int32_t flags, void *stream);
/* matmulGrad (PISOtorch.cpp:613-621; matrix_vector_ops_grads.cu:217-310): a_grad, b_grad
   (written completely) from out_grad.  An inverted operand's gradient is zero; with
   matrix x matrix, both are zero when the output or both operands are inverted. */
pict_status pict_matmul_grad(const void *a, int32_t a_is_matrix, const void *b, int32_t b_is_matrix,
                             const void *out_grad, void *a_grad, void *b_grad, int32_t num_dims, int64_t cells,
                             int32_t dtype, int32_t flags, void *stream);
/* VectorToDiagMatrix (PISOtorch.cpp:622-623; matrix_vector_ops.cu:203-262): matrices [C = D^2]
   with the vector on the diagonal and zeros elsewhere (written completely). */
pict_status pict_vector_to_diag_matrix(const void *vectors, void *matrices, int32_t num_dims, int64_t cells,
                                       int32_t dtype, void *stream);
/* InvertMatrix (PISOtorch.cpp:624-625; matrix_vector_ops.cu:265-303): out = inverse per cell
   (adjugate / determinant); out may be `matrices` (in place). */
pict_status pict_invert_matrix(const void *matrices, void *out, int32_t num_dims, int64_t cells, int32_t dtype,
                               void *stream);

/* ------------------------------------------------------------------------ */
/* metrics                                                                   */
/* ------------------------------------------------------------------------ */

/* CoordsToTransforms (grid_gen.cu:284-381): coords [1,D,(Z+1),Y+1,X+1] ->
   out [1,(Z),Y,X,2D^2+1].  vertex_size = {X+1, Y+1, Z+1} (unused axes 1). */
pict_status pict_coords_to_transforms(const void *coords, int32_t num_dims,
                                      const int32_t vertex_size[3], int32_t dtype,
                                      void *out, void *stream);
/* CoordsToFaceTransforms (grid_gen.cu:383-514): out [1,D,(Z+1),Y+1,X+1,2D^2+1]. */
pict_status pict_coords_to_face_transforms(const void *coords, int32_t num_dims,
                                           const int32_t vertex_size[3], int32_t dtype,
                                           void *out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* PICT_HIP_H */
