/*
 * owtk_hip.h - C ABI of libowtk_hip.so, the gfx950 (MI355X) execution policy
 * for the OWT-Krylov block-CSR solve path.
 *
 * Conventions
 *   - Every function returns int: OWTK_OK (0) or a negative owtk_status.
 *     owtk_last_error() returns the message of the calling thread's last
 *     failure.  The header-only C++ layer (owtk/hip.hpp) turns
 *     OWTK_ERROR_INVALID_ARGUMENT into std::invalid_argument and
 *     OWTK_ERROR_ZERO_PIVOT into std::runtime_error("zero pivot in ILU(0)"),
 *     the exception types the reference throws (block_csr.hpp:281-300,
 *     preconditioner.hpp:451,480).
 *   - Host arrays passed to *_create / *_update_values are borrowed for the
 *     duration of the call only; the library keeps device copies.
 *   - Device vectors are plain device pointers in the reference BlockVector
 *     layout: [local node][component], owned nodes first, then ghosts
 *     (core.hpp:193-196).  Vector element type follows the operation's dtype:
 *     OWTK_F64 -> double, OWTK_F32 and OWTK_BF16 (matrix storage) -> float.
 *   - Indices are int32 (row offsets and columns).
 *   - Objects are bound to the context that created them.  All work of a
 *     context runs on its one HIP stream, in call order.  Calls on one
 *     context are not reentrant.
 */
#ifndef OWTK_HIP_H
#define OWTK_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status -- */
typedef enum owtk_status {
    OWTK_OK = 0,
    OWTK_ERROR_INVALID_ARGUMENT = -1, /* std::invalid_argument (layout/storage) */
    OWTK_ERROR_ZERO_PIVOT = -2,       /* std::runtime_error("zero pivot in ILU(0)") */
    OWTK_ERROR_HIP = -3,              /* HIP runtime failure */
    OWTK_ERROR_NO_DEVICE = -4,        /* no usable gfx950 device */
    OWTK_ERROR_OUT_OF_MEMORY = -5,
    OWTK_ERROR_UNSUPPORTED = -6,      /* option outside the implemented path */
    OWTK_ERROR_RUNTIME = -7           /* other std::runtime_error */
} owtk_status;

typedef enum owtk_dtype {
    OWTK_F64 = 0,
    OWTK_F32 = 1,
    OWTK_BF16 = 2
} owtk_dtype;

/* Message of the calling thread's last failed call ("" if none). */
const char* owtk_last_error(void);
/* Library version string. */
const char* owtk_version(void);

/* --------------------------------------------------------------- context -- */
typedef struct owtk_context_s* owtk_context;

/* Number of visible HIP devices (0 on a machine without a GPU). */
int owtk_device_count(int* count);
/* Create a context (one HIP stream) on `device`.  Replaces the per-call data
 * mapping of OpenMPTargetExecutionPolicy (execution.hpp:36-90) with
 * device-resident state. */
int owtk_context_create(int device, owtk_context* out);
int owtk_context_destroy(owtk_context ctx);
int owtk_context_synchronize(owtk_context ctx);
/* Name of the policy ("hip-gfx950"), cf. HostExecutionPolicy::name()
 * (execution.hpp:26). */
const char* owtk_policy_name(void);

/* Live kernel timing with HIP events on the context stream; plays the role of
 * the reference's ScopedSolverTimer / TimedOperator instrumentation
 * (core.hpp:150-182, benchmarks/benchmark_split.cpp:107-121) per kernel class.
 * Categories: 0 SpMV FP64, 1 SpMV FP32, 2 SpMV BF16, 3 preconditioner apply
 * FP64, 4 preconditioner apply FP32/BF16, 5 learned preconditioner series
 * steps (middle), 6 learned preconditioner first and last steps. */
int owtk_profile_enable(owtk_context ctx, int enable);
int owtk_profile_reset(owtk_context ctx);
/* Sum over recorded launches of one category: launch count, total kernel
 * milliseconds, total algorithmic HBM bytes. */
/* Region timing: one HIP event pair on the context stream around whatever is
 * launched between begin (begin = 1) and end (begin = 0, returns ms). */
int owtk_profile_region(owtk_context ctx, int begin, double* ms);
int owtk_profile_read(owtk_context ctx, int category, uint64_t* launches, double* total_ms,
                      int64_t* bytes);

/* ---------------------------------------------------------- device memory -- */
int owtk_malloc(owtk_context ctx, size_t bytes, void** device_pointer);
int owtk_free(owtk_context ctx, void* device_pointer);
int owtk_memcpy_htod(owtk_context ctx, void* device_dst, const void* host_src, size_t bytes);
int owtk_memcpy_dtoh(owtk_context ctx, void* host_dst, const void* device_src, size_t bytes);
int owtk_memcpy_dtod(owtk_context ctx, void* device_dst, const void* device_src, size_t bytes);
/* Zero `bytes` bytes with a fill kernel (safe inside captured graphs). */
int owtk_fill_zero(owtk_context ctx, void* device_pointer, size_t bytes);
/* Element conversion between vector types (double <-> float). */
int owtk_convert(owtk_context ctx, size_t count, owtk_dtype dst_type, void* dst,
                 owtk_dtype src_type, const void* src);

/* -------------------------------------------------------------- matrices -- */
typedef struct owtk_matrix_s* owtk_matrix;

typedef enum owtk_matrix_kind {
    OWTK_MATRIX_CSR = 0,   /* BlockCsrMatrix: component-diagonal [entry][comp] (block_csr.hpp:17-21) */
    OWTK_MATRIX_SPLIT = 1, /* SplitBlockCsrMatrixView: diag [node][comp] + edges [edge][comp] (split_block_csr.hpp:17-28) */
    OWTK_MATRIX_BSR = 2    /* DenseBlockCsrMatrix: row-major bs x bs blocks (dense_block_csr.hpp:15-21) */
} owtk_matrix_kind;

typedef struct owtk_matrix_info {
    int32_t kind;
    int32_t block_size;
    int64_t owned_nodes;
    int64_t ghost_nodes;
    int64_t entries;            /* stored block entries (split: off-diagonal only) */
    int64_t device_bytes;       /* device bytes held by the matrix (all storages) */
    int64_t spmv_bytes_f64;     /* algorithmic HBM bytes of one FP64 SpMV */
    int64_t spmv_bytes_f32;     /* ... with FP32 values, FP32 vectors */
    int64_t spmv_bytes_bf16;    /* ... with BF16 values, FP32 vectors */
    int32_t has_f32;            /* low-precision shadows present */
    int32_t has_bf16;
} owtk_matrix_info;

/* BlockCsrMatrix(owned, ghost, block, offsets, columns, values)
 * (block_csr.hpp:27-43).  values: entries * block doubles. */
int owtk_matrix_create_csr(owtk_context ctx, int64_t owned_nodes, int64_t ghost_nodes,
                           int64_t block_size, const int32_t* row_offsets,
                           const int32_t* columns, const double* values,
                           owtk_matrix* out);
/* SplitBlockCsrMatrixView(owned, ghost, block, offsets, columns, diagonal,
 * off_diagonal) (split_block_csr.hpp:32-49). */
int owtk_matrix_create_split(owtk_context ctx, int64_t owned_nodes, int64_t ghost_nodes,
                             int64_t block_size, const int32_t* off_diagonal_row_offsets,
                             const int32_t* off_diagonal_columns, const double* diagonal,
                             const double* off_diagonal, owtk_matrix* out);
/* DenseBlockCsrMatrix(owned, ghost, block, offsets, columns, values)
 * (dense_block_csr.hpp:26-42).  values: entries * block * block doubles. */
int owtk_matrix_create_bsr(owtk_context ctx, int64_t owned_nodes, int64_t ghost_nodes,
                           int64_t block_size, const int32_t* row_offsets,
                           const int32_t* columns, const double* values,
                           owtk_matrix* out);
/* Numeric update with an unchanged structure.  The reference split view
 * borrows its value arrays, so an application's edit is visible at the next
 * apply (split_block_csr.hpp:25-28); on the device that visibility is this
 * explicit call.  CSR/BSR: `values` only (second pointer NULL).  Split:
 * `values` = diagonal, `off_diagonal` = edges.  Low-precision shadows are
 * refreshed too. */
int owtk_matrix_update_values(owtk_matrix m, const double* values, const double* off_diagonal);
/* Add (or refresh) an FP32 or BF16 copy of the values, used by dtype-selected
 * SpMVs and by the mixed-precision solvers. */
int owtk_matrix_add_storage(owtk_matrix m, owtk_dtype dtype);
int owtk_matrix_info_get(owtk_matrix m, owtk_matrix_info* info);
int owtk_matrix_destroy(owtk_matrix m);

/* y = A x (apply / apply_rows over all owned rows: block_csr.hpp:192-215,
 * split_block_csr.hpp:89-119, dense_block_csr.hpp:173-205).  dtype selects the
 * value storage; x has (owned+ghost)*block elements, y owned*block. */
int owtk_spmv(owtk_matrix m, owtk_dtype dtype, const void* x, void* y);
/* r = b - A x (detail::true_residual, krylov_solvers.hpp:62-74), FP64. */
int owtk_residual(owtk_matrix m, const double* b, const double* x, double* r);

/* -------------------------------------------------------- preconditioners -- */
typedef struct owtk_precond_s* owtk_precond;

typedef enum owtk_precond_kind {
    OWTK_PC_IDENTITY = 0, /* IdentityPreconditioner (preconditioner.hpp:17-23) */
    OWTK_PC_JACOBI = 1,   /* JacobiPreconditioner (preconditioner.hpp:25-117) */
    OWTK_PC_SSOR = 2,     /* LocalSsorPreconditioner / SplitLocalSsorPreconditioner
                             (preconditioner.hpp:124-247, split_block_csr.hpp:204-332);
                             dense BSR: block SSOR (no reference counterpart) */
    OWTK_PC_ILU = 3,      /* IluLevelPreconditioner / Ilu0Preconditioner
                             (preconditioner.hpp:253-508) */
    OWTK_PC_LEARNED = 4   /* vertex-block graph network (no reference counterpart) */
} owtk_precond_kind;

typedef struct owtk_precond_params {
    int32_t kind;             /* owtk_precond_kind */
    int32_t storage;          /* owtk_dtype of the preconditioner's values */
    double omega;             /* SSOR relaxation, (0, 2) */
    double minimum_diagonal;  /* Jacobi singular-diagonal bound; <= 0: 64*eps */
    int64_t block_nodes;      /* block-Jacobi partition: consecutive node ranges
                                 of this size; 0 = one block over all owned nodes */
    int32_t fill_level;       /* ILU(k) level (0 = ILU(0)) */
    int32_t schedule;         /* triangular sweeps: 0 auto, 1 level sets, 2 one-wave chains
                                 (same results; a performance choice) */
    void* learned_model;      /* owtk_gnn for OWTK_PC_LEARNED */
} owtk_precond_params;

void owtk_precond_params_default(owtk_precond_params* params, int32_t kind);
int owtk_precond_create(owtk_matrix m, const owtk_precond_params* params, owtk_precond* out);
/* Numeric refresh from the matrix's current device values, reusing all
 * symbolic work (update_values: preconditioner.hpp:55-71, 142-174, 269-311). */
int owtk_precond_update_values(owtk_precond pc);
/* z = M^{-1} r; in == out is allowed (preconditioner.hpp:332-333).  Vector
 * type: double for F64 storage, float otherwise. */
int owtk_precond_apply(owtk_precond pc, const void* in, void* out);
int owtk_precond_numeric_updates(owtk_precond pc, uint64_t* count);
typedef struct owtk_precond_info {
    int32_t kind;
    int32_t storage;
    int64_t blocks;
    int64_t forward_levels;   /* longest forward level schedule over all blocks */
    int64_t backward_levels;
    int64_t factor_entries;   /* stored factor entries (ILU) */
    int64_t apply_bytes;      /* algorithmic HBM bytes of one apply */
    int64_t rows;             /* owned vertices */
    int64_t block_size;
} owtk_precond_info;
int owtk_precond_info_get(owtk_precond pc, owtk_precond_info* info);
/* Copy the ILU factors to the host in the reference layout (factor_row_offsets,
 * factor_columns, factor_values, inverse_diagonal: preconditioner.hpp:365-370).
 * Pass NULL arrays to query sizes through *entries. */
int owtk_precond_ilu_factors(owtk_precond pc, int64_t* entries, int64_t* row_offsets,
                             int64_t* columns, double* values, double* inverse_diagonal);
int owtk_precond_destroy(owtk_precond pc);

/* ---------------------------------------------------------------- solvers -- */
/* How a solve reduces its inner products and norms: the device counterparts
 * of the reference's Reduction template argument (reduction.hpp:103-299).
 * SERIAL: FP64 sums over a fixed tree (SerialReduction).  COMPENSATED: every
 * product and partial sum carried as a double-double pair (TwoProd/TwoSum at
 * every addition, in the threads, the block trees and the finisher), the
 * counterpart of CompensatedSerialReduction's Neumaier sums.
 * MIXED_PRECISION: MixedPrecisionSerialReduction, whose local sums are wider
 * than the vectors: for FP64 vectors the same double-double accumulation (at
 * least the 64-bit significand of x86 long double), for FP32 vectors FP64. */
typedef enum owtk_reduction {
    OWTK_REDUCTION_SERIAL = 0,
    OWTK_REDUCTION_COMPENSATED = 1,
    OWTK_REDUCTION_MIXED_PRECISION = 2
} owtk_reduction;

/* SolverOptions::convergence_test (core.hpp:100-108): called with the
 * iteration count and the current solution (a DEVICE pointer to `count`
 * owned values, FP64) at iteration zero and after every complete iterate;
 * a non-zero return stops the solve, which then measures the true residual
 * and reports converged with converged_by_application set.  While a test is
 * set, the residual tolerance is not tested (the reference's
 * convergence_threshold returns 0).  The solution must not be modified; the
 * test runs with the context's stream idle and may read the solution with
 * owtk_memcpy_dtoh on the same context. */
typedef int (*owtk_convergence_test)(void* user_data, uint64_t iteration, const double* device_solution,
                                     size_t count);

/* Mirrors SolverOptions<T> (core.hpp:88-111); enum values mirror
 * GmresOrthogonalization (core.hpp:32-39). */
typedef struct owtk_solver_options {
    double relative_tolerance;
    double absolute_tolerance;
    double breakdown_tolerance;
    uint64_t maximum_iterations;
    uint64_t convergence_check_interval;
    uint64_t restart;
    uint64_t residual_replacement_interval;
    int32_t gmres_orthogonalization; /* 0 iterated CGS, 1 adaptive CGS, 2 one-sync CGS */
    int32_t collect_timings;
    uint64_t idr_shadow_space;       /* IDR(s): s (SolverOptions::idr_shadow_space, core.hpp:97) */
    double relaxation;               /* SolverOptions::relaxation (core.hpp:98): the stationary
                                        solvers' weight; the Krylov solvers do not read it */
    int32_t verify_true_residual;    /* SolverOptions::verify_true_residual (core.hpp:99): the
                                        device solvers always verify the true residual before
                                        reporting convergence, as the reference's do */
    int32_t reduction;               /* owtk_reduction; non-SERIAL: owtk_bicgstab only */
    owtk_convergence_test convergence_test;  /* NULL: tolerance test (owtk_bicgstab only) */
    void* convergence_test_data;
} owtk_solver_options;

/* Mirrors SolverResult<T> (core.hpp:118-146); status and breakdown_reason use
 * the numeric values of SolverStatus (core.hpp:24-30) and BreakdownReason
 * (core.hpp:41-52). */
typedef struct owtk_solver_result {
    int32_t status;
    int32_t breakdown_reason;
    uint64_t iterations;
    uint64_t operator_applications;
    uint64_t preconditioner_applications;
    uint64_t global_reductions;
    uint64_t reorthogonalizations;
    uint64_t arnoldi_norm_verifications;
    double initial_residual_norm;
    double recursive_residual_norm;
    double true_residual_norm;
    double relative_residual_norm;
    double solve_seconds;
    int32_t converged_by_application; /* SolverResult::converged_by_application (core.hpp:127) */
    /* iterative refinement (owtk_fgmres_refined) only */
    uint64_t refinement_steps;
    uint64_t inner_iterations;
    uint64_t fallback_refinements; /* refinement steps run with the fallback preconditioner */
} owtk_solver_result;

/* Defaults of SolverOptions<double>. */
void owtk_solver_options_default(owtk_solver_options* options);

/* bicgstab(op, b, x, options, pc, reduction) (krylov_solvers.hpp:542-745) with
 * every vector resident on the device.  pc == NULL means
 * IdentityPreconditioner.  options->reduction selects the reduction
 * (bicgstab_stable and bicgstab_mixed_precision, krylov_solvers.hpp:766-809,
 * are this call with OWTK_REDUCTION_MIXED_PRECISION); options->convergence_test
 * is honoured. */
int owtk_bicgstab(owtk_matrix m, owtk_precond pc, const double* b, double* x,
                  const owtk_solver_options* options, owtk_solver_result* result);
/* pipelined_bicgstab(op, b, x, options, pc) (pipelined_bicgstab.hpp:20-244):
 * the fused-reduction variant (three global reductions per iteration), the
 * recursive residual tested every iteration, replacement every
 * residual_replacement_interval iterations restarting the recurrence. */
int owtk_pipelined_bicgstab(owtk_matrix m, owtk_precond pc, const double* b, double* x,
                            const owtk_solver_options* options, owtk_solver_result* result);
/* communication_hiding_bicgstab(op, b, x, options, pc)
 * (pipelined_bicgstab.hpp:317-597): the 15-vector Cools/PETSc recurrence,
 * two fused reductions per iteration, each followed by one preconditioner +
 * operator application; false convergence and replacement restart the
 * recurrence (a failed restart is a biorthogonality_loss breakdown). */
int owtk_communication_hiding_bicgstab(owtk_matrix m, owtk_precond pc, const double* b, double* x,
                                       const owtk_solver_options* options, owtk_solver_result* result);
/* idrs(op, b, x, options, pc) (idrs.hpp:128-505): right-preconditioned IDR(s)
 * with s = options->idr_shadow_space (2..16 on the device; 1 is routed through
 * bicgstab as in the reference).  Shadow vectors are the reference's draws
 * (std::mt19937_64 seeded 42 + restart * 0x9e3779b97f4a7c15, normal
 * distribution), orthonormalised on the device; the moment matrix and small
 * dense solves run on the host.  Counters follow the reference's: one
 * global reduction per inner product or norm it takes. */
int owtk_idrs(owtk_matrix m, owtk_precond pc, const double* b, double* x,
              const owtk_solver_options* options, owtk_solver_result* result);
/* fgmres(op, b, x, options, pc) (krylov_solvers.hpp:159-486), device-resident.
 * An application convergence test (options->convergence_test) sees every
 * iterate x + Z y, formed without touching x (krylov_solvers.hpp:398-427);
 * it costs one host round trip per Arnoldi column.  Serial reduction only. */
int owtk_fgmres(owtk_matrix m, owtk_precond pc, const double* b, double* x,
                const owtk_solver_options* options, owtk_solver_result* result);

/* ArnoldiSnapshot<double> (core.hpp:462-534): the last restart cycle's
 * unrotated Hessenberg columns and preconditioned basis vectors Z, as the
 * fgmres(..., workspace, arnoldi_snapshot) argument fills it
 * (krylov_solvers.hpp:171, 252-253, 361-366, 420-436).  The basis vectors
 * stay on the device in the solve's owned + ghost layout, ghosts zero. */
typedef struct owtk_arnoldi_snapshot_s* owtk_arnoldi_snapshot;
int owtk_arnoldi_snapshot_create(owtk_context ctx, owtk_arnoldi_snapshot* out);
int owtk_arnoldi_snapshot_destroy(owtk_arnoldi_snapshot snapshot);
/* ArnoldiSnapshot::clear (core.hpp:504-510) */
int owtk_arnoldi_snapshot_clear(owtk_arnoldi_snapshot snapshot);
/* ArnoldiSnapshot::columns (core.hpp:512) */
int owtk_arnoldi_snapshot_columns(owtk_arnoldi_snapshot snapshot, uint64_t* columns);
/* ArnoldiSnapshot::h(row, column) (core.hpp:513-519): INVALID_ARGUMENT where
 * the reference throws std::out_of_range (column >= columns or row > columns). */
int owtk_arnoldi_snapshot_h(owtk_arnoldi_snapshot snapshot, uint64_t row, uint64_t column, double* value);
/* ArnoldiSnapshot::preconditioned_basis(column) (core.hpp:520-527): the
 * device address of Z_column (owned + ghost values), valid until the next
 * solve with this snapshot or its destruction. */
int owtk_arnoldi_snapshot_basis(owtk_arnoldi_snapshot snapshot, uint64_t column, const double** device_vector);
/* fgmres(op, b, x, options, pc, SerialReduction, nullptr, arnoldi_snapshot)
 * (krylov_solvers.hpp:159-171): owtk_fgmres that also records the last
 * cycle into `snapshot` (null: owtk_fgmres). */
int owtk_fgmres_with_snapshot(owtk_matrix m, owtk_precond pc, const double* b, double* x,
                              const owtk_solver_options* options, owtk_arnoldi_snapshot snapshot,
                              owtk_solver_result* result);

/* Mixed precision: FP64 iterative refinement around an inner FGMRES whose
 * matrix and preconditioner values are stored in `inner_storage` (F32/BF16)
 * with FP32 vectors and arithmetic; outer residual, update, and all inner
 * products and norms in FP64.  No reference counterpart. */
typedef struct owtk_refinement_options {
    owtk_solver_options outer;   /* relative/absolute tolerance and limits of the FP64 loop */
    int32_t inner_storage;       /* OWTK_F32 or OWTK_BF16 */
    int32_t inner_restart;       /* inner FGMRES(m) */
    double inner_relative_tolerance;
    uint64_t inner_maximum_iterations;
    uint64_t maximum_refinements;
    double fallback_ratio;       /* owtk_fgmres_refined_fallback: switch when an outer step
                                    leaves more than this fraction of the true residual (0.5) */
} owtk_refinement_options;
void owtk_refinement_options_default(owtk_refinement_options* options);
int owtk_fgmres_refined(owtk_matrix m, owtk_precond inner_pc, const double* b, double* x,
                        const owtk_refinement_options* options, owtk_solver_result* result);
/* The same driver with a fallback inner preconditioner (no reference
 * counterpart): the true FP64 residual judges every outer step; when a step
 * with inner_pc leaves more than fallback_ratio of the previous true residual
 * (or the inner solve breaks down or produces non-finite values), the step is
 * undone and the remaining steps use fallback_pc (e.g. learned -> block-Jacobi
 * ILU(0)).  result->fallback_refinements counts the steps that used it. */
int owtk_fgmres_refined_fallback(owtk_matrix m, owtk_precond inner_pc, owtk_precond fallback_pc,
                                 const double* b, double* x, const owtk_refinement_options* options,
                                 owtk_solver_result* result);

/* ------------------------------------------------------------- recycling -- */
/* RecycleSpace<double> (recycling.hpp:24-187): up to `capacity` pairs
 * (U_i, C_i = A U_i) with C orthonormal, kept on the device across solves;
 * `layout` fixes the vector layout (owned + ghost entries).  All vector
 * arguments are device pointers.  telemetry (nullable) receives the
 * operator applications and global reductions, as SolverResult* telemetry
 * in the reference. */
typedef struct owtk_recycle_space_s* owtk_recycle_space;
int owtk_recycle_space_create(owtk_matrix layout, uint64_t capacity, owtk_recycle_space* out);
int owtk_recycle_space_destroy(owtk_recycle_space space);
int owtk_recycle_space_size(owtk_recycle_space space, int64_t* size);
int owtk_recycle_space_clear(owtk_recycle_space space);
/* copies of U_i and C_i (owned entries) */
int owtk_recycle_space_get(owtk_recycle_space space, int64_t index, double* u, double* c);
/* add_candidate (recycling.hpp:79-135); *added = 0 when the candidate is
 * numerically contained in the space */
int owtk_recycle_space_add_candidate(owtk_recycle_space space, owtk_matrix op, const double* candidate,
                                     double dependence_tolerance, int32_t* added, owtk_solver_result* telemetry);
/* refresh (recycling.hpp:65-73): C = A U after an operator update */
int owtk_recycle_space_refresh(owtk_recycle_space space, owtk_matrix op, owtk_solver_result* telemetry);
/* project_initial_guess (recycling.hpp:138-163): x += U C^T residual */
int owtk_recycle_space_project(owtk_recycle_space space, const double* residual, double* x,
                               owtk_solver_result* telemetry);
/* recycled_fgmres (recycling.hpp:291-360): augmented FGMRES; with
 * update_recycle_space the converged correction x - x0 becomes the next
 * candidate */
int owtk_recycled_fgmres(owtk_matrix m, owtk_precond pc, const double* b, double* x, owtk_recycle_space space,
                         const owtk_solver_options* options, int32_t update_recycle_space,
                         owtk_solver_result* result);
/* recycled_fgmres with its trailing ArnoldiSnapshot* argument
 * (recycling.hpp:300-310): the last FGMRES cycle is recorded into `snapshot`
 * (may be null) */
int owtk_recycled_fgmres_with_snapshot(owtk_matrix m, owtk_precond pc, const double* b, double* x,
                                       owtk_recycle_space space, const owtk_solver_options* options,
                                       int32_t update_recycle_space, owtk_arnoldi_snapshot snapshot,
                                       owtk_solver_result* result);

/* ---------------------------------------------------------------- GCRO-DR -- */
/* ArnoldiSnapshot::begin_cycle / record_column / finish_cycle
 * (core.hpp:465-500), for snapshots an application assembles itself:
 * begin_cycle takes the vector layout of `layout` (owned + ghost of that
 * matrix); record_column takes column+2 host values; finish_cycle copies the
 * owned values of `columns` device vectors. */
int owtk_arnoldi_snapshot_begin_cycle(owtk_arnoldi_snapshot snapshot, owtk_matrix layout,
                                      uint64_t maximum_columns);
int owtk_arnoldi_snapshot_record_column(owtk_arnoldi_snapshot snapshot, uint64_t column, const double* values,
                                        size_t count);
int owtk_arnoldi_snapshot_finish_cycle(owtk_arnoldi_snapshot snapshot, const double* const* basis,
                                       uint64_t columns);
/* The harmonic-Ritz eigenproblem harmonic y = theta metric y (n x n,
 * column-major; harmonic symmetric positive definite) with the output of
 * LAPACK dggev as gcrodr.hpp:30-78 calls it: theta_j = (real_j + i imag_j)
 * / beta_j (beta_j = 0: infinite), complex pairs consecutive with imag > 0
 * first and their right eigenvector as (real, imaginary) columns of `right`
 * (n x n).  Host computation: Cholesky of `harmonic`, then Hessenberg QR and
 * inverse iteration on the equivalent standard problem. */
int owtk_harmonic_ggev(int n, const double* harmonic, const double* metric, double* real, double* imag,
                       double* beta, double* right);
/* extract_harmonic_ritz(op, snapshot, space, reduction, telemetry)
 * (gcrodr.hpp:86-181): replace `space` with the harmonic Ritz vectors of the
 * snapshot's basis Z for A = m (smallest |theta| first, a complex pair only
 * when both directions fit), each added with add_candidate.  *size receives
 * the new size; telemetry (may be null) is incremented. */
int owtk_extract_harmonic_ritz(owtk_matrix m, owtk_arnoldi_snapshot snapshot, owtk_recycle_space space,
                               owtk_solver_result* telemetry, uint64_t* size);
/* gcrodr(op, b, x, space, options, pc, reduction, workspace, snapshot)
 * (gcrodr.hpp:185-212): recycled_fgmres without the correction update,
 * recording the last cycle, then extract_harmonic_ritz.  `snapshot` may be
 * null (a temporary one is used). */
int owtk_gcrodr(owtk_matrix m, owtk_precond pc, const double* b, double* x, owtk_recycle_space space,
                const owtk_solver_options* options, owtk_arnoldi_snapshot snapshot, owtk_solver_result* result);

/* -------------------------------------------------------------- multigrid -- */
/* AggregationMultigrid (multigrid.hpp:41-421): two-level aggregation solver
 * over a component-diagonal BlockCsrMatrix (OWTK_MATRIX_CSR).  Set-up on the
 * host as the reference (greedy aggregates, row-averaged coarse matrix,
 * inverse diagonals, greedy red/black colouring); every cycle operation on
 * the device.  The operator of solve() is the fine matrix itself. */
typedef struct owtk_multigrid_s* owtk_multigrid;
typedef enum owtk_multigrid_smoother {  /* MultigridSmoother (multigrid.hpp:23-27) */
    OWTK_MG_JACOBI = 0,
    OWTK_MG_RED_BLACK = 1,  /* red_black_gauss_seidel */
    OWTK_MG_CHEBYSHEV = 2
} owtk_multigrid_smoother;
typedef enum owtk_multigrid_cycle {     /* MultigridCycle (multigrid.hpp:17-21) */
    OWTK_MG_V = 0,
    OWTK_MG_W = 1,
    OWTK_MG_FULL = 2
} owtk_multigrid_cycle;
typedef struct owtk_multigrid_options { /* MultigridOptions (multigrid.hpp:29-37) */
    uint64_t pre_smoothing_steps;       /* 2 */
    uint64_t post_smoothing_steps;      /* 2 */
    uint64_t coarse_smoothing_steps;    /* 24 */
    uint64_t target_aggregate_size;     /* 4 */
    double jacobi_weight;               /* 0.8 */
    int32_t smoother;                   /* owtk_multigrid_smoother, OWTK_MG_JACOBI */
    uint64_t chebyshev_levels;          /* 8 */
} owtk_multigrid_options;
void owtk_multigrid_options_default(owtk_multigrid_options* options);
/* AggregationMultigrid(fine_matrix, options) */
int owtk_multigrid_create(owtk_matrix fine, const owtk_multigrid_options* options, owtk_multigrid* out);
int owtk_multigrid_destroy(owtk_multigrid mg);
/* coarse_matrix() and fine_to_coarse(): NULL arrays for the sizes first */
int owtk_multigrid_coarse(owtk_multigrid mg, int64_t* nodes, int64_t* entries, int32_t* row_offsets,
                          int32_t* columns, double* values, int64_t* fine_to_coarse);
/* solve(fine_operator, rhs, solution, options, cycle) (multigrid.hpp:74-221) */
int owtk_multigrid_solve(owtk_multigrid mg, const double* b, double* x, const owtk_solver_options* options,
                         int32_t cycle, owtk_solver_result* result);

/* ------------------------------------------------------------ matrix files -- */
/* The reference's binary fixture format (benchmarks/matrix_analysis/
 * ilu_fill.cpp:20-44), little-endian: u64 {n, bins, nnz}, u64 offsets[n+1],
 * u64 columns[nnz], f64 values[nnz*bins] (component-diagonal BlockCsrMatrix,
 * diagonal included).  Read: call with NULL arrays for the sizes, then with
 * arrays of n+1 / nnz / nnz*bins elements (int32 indices on the device path). */
int owtk_read_block_csr(const char* path, int64_t* nodes, int64_t* block_size, int64_t* entries,
                        int32_t* row_offsets, int32_t* columns, double* values);
int owtk_write_block_csr(const char* path, int64_t nodes, int64_t block_size, const int32_t* row_offsets,
                         const int32_t* columns, const double* values);
/* ILU factors in the format ilu_fill.cpp writes (ilu_fill.cpp:36-41): u64
 * {n, bins, factor_nnz}, u64 factor offsets, u64 factor columns, f64 factor
 * values, f64 inverse diagonal. */
int owtk_write_ilu_factors(owtk_precond pc, const char* path);
/* read_state (benchmarks/specwave_limon.py:66-100): the application's per-rank
 * final state files, "SWSTATE1" + u64 {width, owned, ns, nd, count} + u64
 * ids[owned] + f32/f64 values[owned][ns*nd].  The files together must cover
 * global vertices 0..count-1 exactly once with finite, non-negative values;
 * the same malformed inputs are rejected.  values == NULL: *vertices only;
 * else values receives count*ns*nd doubles ordered by global id. */
int owtk_read_state(const char* const* paths, int32_t files, int64_t ns, int64_t nd, int64_t* vertices,
                    double* values);

/* ------------------------------------------- learned preconditioner (GNN) -- */
/* No reference counterpart: the reference defers learned preconditioners
 * (REVIEW_AND_IMPROVEMENT_PLAN.md:377-379).  A vertex-block graph network
 * (3 message-passing layers, width 32, BF16 weights, FP32 arithmetic) maps a
 * split matrix to per-vertex, per-component weights gamma_k, k = 0..degree,
 * of the preconditioner
 *     u_0 = D^-1 r,  u_(k+1) = -D^-1 A_off u_k,  z = sum_k gamma_k u_k
 * (OWTK_PC_LEARNED with params.learned_model = the model; storage selects the
 * precision of the preconditioner's own copy of the off-diagonal values,
 * built from the FP64 values at set-up and at owtk_precond_update_values).
 * The apply evaluates the same series through w_k = D u_k over column-scaled
 * values, so no step reads D^-1.  Network: features per vertex and component
 * rho = sum|a|/d, log(1 - rho)/8 and the edge anisotropy
 * log(deg max|a| / sum|a|)/2, plus deg/8; decoder gamma_kic = g_kc + o_kci
 * (o = W_out h3 + b_out).  A fresh model has Glorot-uniform encoder and
 * layer weights (a fixed seed), a zero decoder and g = 1: the plain
 * truncated Neumann series.
 * Training: each problem holds a split matrix and P (residual, error) pairs;
 * the loss is (1/P) sum_p |D^-1/2 A (M r_p - e_p)|^2 / |D^-1/2 A e_p|^2.
 * Gradients accumulate on the device (owtk_gnn_accumulate) and an Adam step
 * applies and clears them.  Parameters are a flat FP32 vector in the order
 * W_in[32][3b+1], b_in[32], then per layer W_s[32][32], W_n[32][32], b[32],
 * then W_out[(degree+1)b][32], b_out, g[(degree+1)b]. */
typedef struct owtk_gnn_s* owtk_gnn;
typedef struct owtk_gnn_problem_s* owtk_gnn_problem;

int owtk_gnn_create(owtk_context ctx, int32_t block_size, int32_t degree, owtk_gnn* out);
int owtk_gnn_destroy(owtk_gnn model);
int owtk_gnn_info(owtk_gnn model, int32_t* block_size, int32_t* degree, int64_t* parameter_count);
/* Host FP32 parameters in and out; setting them resets the Adam state. */
int owtk_gnn_set_parameters(owtk_gnn model, const float* params);
int owtk_gnn_get_parameters(owtk_gnn model, float* params);
/* Network inference on a split matrix: gamma[k][vertex*b + c] to host. */
int owtk_gnn_coefficients(owtk_gnn model, owtk_matrix m, float* gamma);
/* residuals/errors: device FP64, pairs x (owned*b).  The pattern must be
 * structurally symmetric (the backward pass uses the transposed matrix).
 * Every model of the same context, block size and degree can train on the
 * problem (owtk_gnn_accumulate). */
int owtk_gnn_problem_create(owtk_gnn model, owtk_matrix m, int32_t pairs, const double* residuals,
                            const double* errors, owtk_gnn_problem* out);
int owtk_gnn_problem_destroy(owtk_gnn_problem problem);
int owtk_gnn_zero_grad(owtk_gnn model);
/* gradient += weight * dL/dtheta, accumulated loss += weight * L (device). */
int owtk_gnn_accumulate(owtk_gnn model, owtk_gnn_problem problem, double weight);
/* Accumulated loss and gradient to host (either may be NULL). */
int owtk_gnn_read(owtk_gnn model, double* loss, float* gradient);
/* Adam update from the accumulated gradient; returns the accumulated loss
 * (may be NULL) and clears gradient and loss. */
int owtk_gnn_adam_step(owtk_gnn model, double learning_rate, double beta1, double beta2, double epsilon,
                       double* loss);

/* ------------------------------------------------------- batched BiCGSTAB -- */
/* No reference counterpart (the reference solves one system per call).  Many
 * independent component-diagonal block-CSR systems (BlockCsrMatrix form,
 * block_csr.hpp:192-236, diagonal entry included) of uneven size, solved by
 * BiCGSTAB + Jacobi with the recurrence, breakdown guards, check cadence and
 * residual replacement of bicgstab (krylov_solvers.hpp:542-745) per system,
 * all systems in one hipGraph-captured iteration sequence with per-system
 * convergence masks.  Arrays are the systems' CSR concatenated: system s owns
 * vertices [vertex_offsets[s], vertex_offsets[s+1]); row_offsets (size
 * vertices+1) and columns index the concatenation (columns must stay inside
 * the row's system); values are [entry][component]. */
typedef struct owtk_batch_s* owtk_batch;
typedef struct owtk_batch_result {
    int32_t status;            /* owtk_solver_status */
    int32_t breakdown_reason;
    uint64_t iterations;
    double true_residual_norm; /* last true residual |b - A x| of the system */
    double rhs_norm;
    double relative_residual_norm;
    double stop_ms;            /* ms from the solve's start to the graph replay
                                  in which the system stopped (a workgroup-
                                  solved system: to the device clock's reading
                                  when it stopped) */
} owtk_batch_result;

int owtk_batch_create(owtk_context ctx, int32_t systems, int32_t block_size, const int64_t* vertex_offsets,
                      const int64_t* row_offsets, const int32_t* columns, const double* values,
                      owtk_batch* out);
int owtk_batch_destroy(owtk_batch batch);
/* Rows per workgroup tile: 64, 128, 256 or 512 (no tile spans two systems). */
int owtk_batch_set_tile(owtk_batch batch, int32_t tile_rows);
/* Times `iterations` iterations with every system active for each tile size
 * and keeps the fastest; returns it and its ms per iteration. */
int owtk_batch_autotune(owtk_batch batch, const double* b, int32_t iterations, int32_t* tile_rows,
                        double* ms_per_iteration);
/* Systems of at most max_vertices vertices are solved whole by one workgroup
 * of `threads` threads (256, 512 or 1024) each, in one launch that loops over
 * the iterations on a second stream while the graph replays run the others;
 * 0 (the default) puts every system on the graph path.  Same recurrence and
 * stages; the reductions are summed in another (fixed) order. */
int owtk_batch_set_group_limit(owtk_batch batch, int64_t max_vertices, int32_t threads);
/* Times whole solves of b (x0 = 0, `options`) for limits 0, 1024, 2048, ...
 * (up to the largest system, at most 65,536) with 1024 threads, then 256 and
 * 512 threads at the best limit, and keeps the fastest; returns its limit,
 * threads and ms per solve. */
int owtk_batch_autotune_groups(owtk_batch batch, const double* b, const owtk_solver_options* options,
                               int32_t graph_iterations, int64_t* max_vertices, int32_t* threads, double* ms);
/* b, x: device FP64 over all systems' unknowns (x holds the initial guesses).
 * graph_iterations: iterations per captured graph (the host checks the
 * number of running systems once per replay).  results: host array of
 * `systems` entries, may be NULL. */
int owtk_batch_bicgstab(owtk_batch batch, const double* b, double* x, const owtk_solver_options* options,
                        int32_t graph_iterations, owtk_batch_result* results);

/* --------------------------------------------------------- vector kernels -- */
/* dot with FP64 accumulation (SerialReduction::dot, reduction.hpp:114-145). */
int owtk_dot(owtk_context ctx, owtk_dtype dtype, size_t n, const void* a, const void* b,
             double* out);
/* overflow-safe 2-norm (detail::scaled_norm, reduction.hpp:64-89). */
int owtk_norm(owtk_context ctx, owtk_dtype dtype, size_t n, const void* a, double* out);
/* Reduction::dot / Reduction::norm with the reduction selected
 * (owtk_reduction: SerialReduction, CompensatedSerialReduction,
 * MixedPrecisionSerialReduction; reduction.hpp:114-168, 176-299). */
int owtk_reduction_dot(owtk_context ctx, int32_t reduction, owtk_dtype dtype, size_t n, const void* a,
                       const void* b, double* out);
int owtk_reduction_norm(owtk_context ctx, int32_t reduction, owtk_dtype dtype, size_t n, const void* a,
                        double* out);
/* y += alpha x (simd::axpy, simd.hpp:110-202). */
int owtk_axpy(owtk_context ctx, size_t n, double alpha, const double* x, double* y);
/* y += alpha x, then out = <w, y> in the same pass. */
int owtk_axpy_dot(owtk_context ctx, size_t n, double alpha, const double* x, double* y,
                  const double* w, double* out);
/* out[j] = <w, v_j> for j < k, out[k] = <w, w>: the batched CGS local dots
 * (krylov_solvers.hpp:273-282). vs: host array of k device pointers. */
int owtk_multidot(owtk_context ctx, size_t n, size_t k, const double* const* vs,
                  const double* w, double* out);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The same multi-dot as owtk_multidot, with bitwise the
 * same results, but it needs no context and allocates nothing: the
 * cross-block partials live in a static device array of 516,096 bytes per
 * device (at most 62 basis vectors, the same limit), held until the process
 * ends.  Runs on the calling thread's current HIP device, on `stream` (a
 * hipStream_t; NULL is the null stream), and returns without waiting.  out:
 * DEVICE array of k+1 doubles.  vs: host array of k device pointers.  Calls
 * that overlap on one device share the array and race. */
int owtk_multidot_noalloc(void* stream, size_t n, size_t k, const double* const* vs,
                          const double* w, double* out);

/* ------------------------------------------- synthetic grids (host only) -- */
typedef enum owtk_ordering {
    OWTK_ORDER_IDENTITY = 0,
    OWTK_ORDER_RCM = 1,      /* reverse_cuthill_mckee_order (ordering.hpp:138-180) */
    OWTK_ORDER_HILBERT = 2   /* hilbert_order (ordering.hpp:231-261) */
} owtk_ordering;

typedef enum owtk_coefficient_law {
    OWTK_LAW_UNIFORM = 0,     /* -U(0.1,1)(1+0.3U(-1,1)) per (edge, component) */
    OWTK_LAW_LOGNORMAL_FIELD = 1, /* times exp(sigma (g_u+g_v)/2), g a correlated unit Gaussian field */
    OWTK_LAW_EDGE_CONTRAST = 2    /* times exp(sigma z_e), z_e iid N(0,1) per undirected edge */
} owtk_coefficient_law;

typedef struct owtk_grid_spec {
    int32_t nx, ny;          /* nodes per direction; quads split along (i,j)-(i+1,j+1) */
    int32_t block_size;
    int32_t ordering;        /* owtk_ordering applied after the seeded permutation */
    uint64_t seed;
    int32_t permute;         /* seeded random node permutation before ordering */
    int32_t coefficient_law; /* owtk_coefficient_law */
    double jitter;           /* node jitter, fraction of h (0.2) */
    double diagonal_shift;   /* diagonal = row |off-diagonal| sum + shift (0.05) */
    double field_sigma;      /* lognormal sigma */
    double field_correlation;/* correlation length in units of h */
} owtk_grid_spec;

void owtk_grid_spec_default(owtk_grid_spec* spec, int32_t nx, int32_t ny, int32_t block_size,
                            uint64_t seed);
/* Node and off-diagonal entry counts of the triangulated grid. */
int owtk_grid_sizes(const owtk_grid_spec* spec, int64_t* nodes, int64_t* off_diagonal_entries);
/* Split layout (also the component-diagonal CSR minus its diagonal).  Any
 * output pointer except offsets/columns/diagonal/off_diagonal may be NULL.
 * coordinates: 2 doubles per node; final_to_natural: natural id j*nx+i. */
int owtk_grid_generate_split(const owtk_grid_spec* spec, int32_t* row_offsets,
                             int32_t* columns, double* diagonal, double* off_diagonal,
                             double* rhs, double* coordinates, int64_t* final_to_natural);
/* Dense BSR: diagonal blocks N(0,1)+6I, off-diagonal blocks N(0,0.15)*U(0.1,1).
 * offsets/columns include the diagonal entry of each row (sorted columns). */
int owtk_grid_generate_bsr(const owtk_grid_spec* spec, int32_t* row_offsets,
                           int32_t* columns, double* values, double* rhs);
/* Merge a split layout into component-diagonal CSR with sorted columns. */
int owtk_split_to_csr(int64_t owned_nodes, int64_t block_size, const int32_t* offsets,
                      const int32_t* columns, const double* diagonal,
                      const double* off_diagonal, int32_t* csr_offsets,
                      int32_t* csr_columns, double* csr_values);
/* Orderings (new_to_old).  rcm: graph of a CSR over owned nodes. */
int owtk_order_rcm(int64_t n, const int32_t* row_offsets, const int32_t* columns,
                   int64_t* new_to_old);
int owtk_order_hilbert(int64_t n, const double* coordinates, int64_t* new_to_old);

#ifdef __cplusplus
}
#endif

#endif /* OWTK_HIP_H */
