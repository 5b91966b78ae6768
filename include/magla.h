/*
 * magla.h -- C ABI of the MI355X (gfx950) UCHOL / CGGP library, libmagla.so.
 *
 * Drop-in boundary for MAGLA's CAESolvers (C#/.NET 8).  A managed wrapper
 * binds these with [LibraryImport] (blittable pointers, opaque handles).
 * Each entry point names the reference member it replaces (paths relative to
 * the MAGLA repository root).
 *
 * Conventions
 *   - The caller owns every input array; the library copies what it keeps.
 *   - Matrices are MAGLA's SymmetricCSRMatrix layout: upper triangle with the
 *     diagonal, rows ascending, columns ascending within a row, int32
 *     indices, float64 values (SymmetricCSRMatrix.cs:373-390).
 *   - Every function returns an mg_status; the message of the last failure
 *     on the calling thread is returned by mg_last_error().
 *   - "_dev" entry points take device pointers (HBM) and a hipStream_t passed
 *     as void*; all others take host pointers and are synchronous.
 */
#ifndef MAGLA_H
#define MAGLA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes; each maps to the exception the reference throws. */
typedef enum {
    MG_OK = 0,
    MG_E_NULL = 1,      /* ArgumentNullException            (SymmetricUtduSolver.cs:72-73,87-88) */
    MG_E_ARG = 2,       /* ArgumentException                (UtduNumericFactorization.cs:96-98, UtduSymbolicFactorization.cs:306-319) */
    MG_E_RANGE = 3,     /* ArgumentOutOfRangeException / IndexOutOfRangeException (IterativeLinearSolver.cs:26-33) */
    MG_E_NOT_PD = 4,    /* InvalidOperationException: non-positive pivot (UtduNumericFactorization.cs:559-566), p'Ap<=0 (ConjugateGradientGaussPreSolver.cs:139-141), zero diagonal (JacobiPreconditioner.cs:27-28) */
    MG_E_TOO_DENSE = 5, /* InvalidOperationException: front/block too large (UtduNumericFactorization.cs:298-304, FactorStorageLayout.cs:73-75) */
    MG_E_DEVICE = 6,    /* InvalidOperationException wrapping a HIP error (UtduNumericFactorization.cs:1017-1022) */
    MG_E_INTERNAL = 7,  /* InvalidOperationException: internal symbolic/AMD error (ApproximateMinimumDegreeOrdering.cs:220-223) */
    MG_E_NOMEM = 8      /* OutOfMemoryException */
} mg_status;

/* new: front-size buckets of the Schur tile autotuner (f < 128, < 256, < 512, < 1024, < 2048, >= 2048) */
#define MG_SCHUR_BUCKETS 6

typedef enum { MG_ORDER_NATURAL = 0, MG_ORDER_AMD = 1 } mg_ordering; /* FillReducingOrdering (UtduSymbolicFactorization.cs:8-24) */
typedef enum { MG_FP64 = 0, MG_FP32 = 1, MG_BF16 = 2 } mg_precision; /* new: factor precision (FP64/FP32);
                                                                         BF16 only for the learned preconditioner */

/* UtduSolverOptions (UtduSymbolicFactorization.cs:31-98) + device fields. */
typedef struct {
    int32_t ordering;                     /* default MG_ORDER_AMD */
    int32_t max_degree_of_parallelism;    /* host threads for the symbolic phase; 0 = all cores */
    int32_t supernode_amalgamation_limit; /* default 64 */
    int32_t panel_width;                  /* default 32, clamped to 4..128 (reference semantics); the
                                             device factorization blocks by 64 independently */
    int32_t allow_diagonal_regularization;/* default 1 */
    int32_t precision;                    /* MG_FP64 (default) or MG_FP32 */
    double pivot_tolerance;               /* default 1e-13 */
    double diagonal_regularization;       /* default 1e-10 */
    int64_t parallel_update_threshold;    /* default 1<<20 (kept for ABI parity; unused on the GPU) */
    int32_t device_graph;                 /* new: 1 = capture the factorization launches in a hipGraph on
                                             the first mg_factorize_dev of a numeric handle and replay it
                                             on refactorization, and capture the handle's solves (first
                                             solve, residual, refinement rounds) and replay them while
                                             their arguments repeat; 2 = the solves only (the
                                             factorization stays eager, with its lookahead stream);
                                             0 = neither (default) */
    int32_t reserved;
} mg_options;

typedef struct mg_symbolic mg_symbolic;   /* UtduSymbolicFactorization */
typedef struct mg_numeric mg_numeric;     /* UtduNumericFactorization */
typedef struct mg_matrix mg_matrix;       /* device-resident SymmetricCSRMatrix (for SpMV / CG / refinement) */

/* UtduSymbolicFactorization public properties (UtduSymbolicFactorization.cs:126-232). */
typedef struct {
    int32_t size;
    int32_t supernode_count;
    int32_t max_front_size;
    int32_t max_supernode_width;
    int64_t factor_entry_count;
    int64_t strict_factor_nonzero_count;
    int64_t factor_operation_count;
    int64_t estimated_factor_bytes;
    int64_t max_front_bytes;
    int64_t peak_contribution_bytes;
    int64_t structure_bytes;
    int32_t level_count;                  /* new: supernodal-tree levels scheduled on the device */
    int32_t matrix_nonzero_count;
} mg_symbolic_info;

/* UtduNumericFactorization statistics (UtduNumericFactorization.cs:57-84). */
typedef struct {
    int32_t regularized_pivot_count;
    int32_t negative_pivot_count;
    double smallest_pivot_magnitude;
    double largest_pivot_magnitude;
    int32_t is_positive_definite;
    int32_t precision;
} mg_numeric_info;

/* Report of a (refined) solve. */
typedef struct {
    int32_t refinement_steps;            /* correction solves after the first one */
    int32_t converged;                   /* max relative residual <= tolerance */
    double relative_residual;            /* max over right-hand sides of ||b-Ax||/||b|| (FP64) */
} mg_solve_report;

/* IterativeSolverResult (CGGP/IterativeSolverResult.cs:8-39), without the solution array. */
typedef struct {
    int32_t iterations;
    int32_t converged;
    double residual_norm;
    double relative_residual;
} mg_pcg_result;

/* ---- library ---- */
const char *mg_last_error(void);
int mg_device_count(int *count);
void mg_options_default(mg_options *opts);   /* new UtduSolverOptions() */

/* ---- host-side matrix assembly (SymmetricCSRMatrix ctor / builder, LineCross) ---- */
/* SymmetricCSRMatrix(int, IEnumerable<MatrixEntry>) (SymmetricCSRMatrix.cs:36-90): (row,col)
 * normalised to (min,max), duplicates summed in input order, |sum|<=1e-15 dropped, columns
 * ascending.  Two-call protocol: pass NULL outputs to get *nnz_out, then call again with
 * rowptr[n+1], colidx[nnz], values[nnz]. */
int mg_csr_from_coo(int32_t n, int64_t count, const int32_t *rows, const int32_t *cols,
                    const double *vals, int32_t *nnz_out, int32_t *rowptr, int32_t *colidx,
                    double *values);
/* SymmetricCSRMatrix.LineCross (SymmetricCSRMatrix.cs:125-167) for a set of constrained
 * equations with prescribed values, applied in the given order; values and rhs updated in place. */
int mg_csr_line_cross(int32_t n, const int32_t *rowptr, const int32_t *colidx, double *values,
                      double *rhs, int32_t count, const int32_t *indices, const double *prescribed);

/* ---- symbolic phase (host) ---- */
/* UtduSymbolicFactorization.Analyze (UtduSymbolicFactorization.cs:240-295) /
 * SymmetricUtduSolver.Analyze (SymmetricUtduSolver.cs:70-76). */
int mg_analyze(int32_t n, const int32_t *rowptr, const int32_t *colidx, const mg_options *opts,
               mg_symbolic **out);
/* Permutation / InversePermutation / cost estimates (UtduSymbolicFactorization.cs:126-232);
 * perm and iperm may be NULL. */
int mg_symbolic_info_get(const mg_symbolic *sym, mg_symbolic_info *info, int32_t *perm, int32_t *iperm);
/* Internal structure for parity tests: EliminationTreeParent, ColumnCounts (n each),
 * fill-reducing order before postorder (n), FirstColumn (nsuper+1), supernode Parent (nsuper),
 * PatternPointers (nsuper+1), PatternRows (pattern_pointers[nsuper]).  Any may be NULL. */
int mg_symbolic_structure(const mg_symbolic *sym, int32_t *etree, int32_t *colcount,
                          int32_t *fill_order, int32_t *first_column, int32_t *super_parent,
                          int32_t *pattern_pointers, int32_t *pattern_rows);
/* UtduSymbolicFactorization.EstimatePeakBytes (UtduSymbolicFactorization.cs:203-219). */
int64_t mg_symbolic_estimate_peak_bytes(const mg_symbolic *sym, int32_t worker_count);
void mg_symbolic_release(mg_symbolic *sym);

/* ---- numeric phase (device) ---- */
/* SymmetricUtduSolver.Factorize(matrix, symbolic) (SymmetricUtduSolver.cs:84-96):
 * EnsureCompatible (n and nnz) then the multifrontal factorization on the GPU. */
int mg_factorize(mg_symbolic *sym, int32_t n, int32_t nnz, const double *values,
                 const mg_options *opts, mg_numeric **out);
/* Same with values already in HBM; refactorizes into *num when *num is non-NULL and was
 * created from the same symbolic handle and precision (no allocation on that path). */
int mg_factorize_dev(mg_symbolic *sym, int32_t nnz, const double *d_values, const mg_options *opts,
                     void *stream, mg_numeric **num);
/* RegularizedPivotCount ... Diagonal (UtduNumericFactorization.cs:57-84); diag may be NULL
 * (length n, elimination order). */
int mg_numeric_info_get(const mg_numeric *num, mg_numeric_info *info, double *diag);
void mg_numeric_release(mg_numeric *num);

/* UtduNumericFactorization.Solve (UtduNumericFactorization.cs:90-112) for nrhs columns
 * (b and x column-major with leading dimensions ldb, ldx >= n).  With an FP32 factor each
 * column is refined in FP64 against the factorized matrix until ||b-Ax||/||b|| <= tol or
 * max_refine correction steps; with an FP64 factor tol/max_refine apply the same way
 * (max_refine = 0: one plain solve).  report may be NULL. */
int mg_solve(mg_numeric *num, int32_t nrhs, const double *b, int32_t ldb, double *x, int32_t ldx,
             double tol, int32_t max_refine, mg_solve_report *report);
int mg_solve_dev(mg_numeric *num, int32_t nrhs, const double *d_b, int32_t ldb, double *d_x,
                 int32_t ldx, double tol, int32_t max_refine, void *stream, mg_solve_report *report);
/* No counterpart in the reference: mg_solve_dev for callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, hosts that forbid allocation).
 * Same arguments, kernels and results; the solve workspace lives in a static 256 MiB device
 * array of the library (one per GPU, held until the process ends) instead of the handle's
 * pool, so no call allocates.  A call whose workspace would exceed 256 MiB (about
 * (2 n + sum of front sizes) x nrhs x factor bytes + 8 n x nrhs x 2) returns MG_E_RANGE.  All
 * calls on one GPU share the array: calls that overlap on one GPU (other threads or streams)
 * race, so callers serialize them. */
int mg_solve_dev_noalloc(mg_numeric *num, int32_t nrhs, const double *d_b, int32_t ldb, double *d_x,
                         int32_t ldx, double tol, int32_t max_refine, void *stream, mg_solve_report *report);

/* ---- device matrix, SpMV, CGGP ---- */
int mg_matrix_create(int32_t n, const int32_t *rowptr, const int32_t *colidx, const double *values,
                     mg_matrix **out);
void mg_matrix_release(mg_matrix *m);
/* SymmetricCSRMatrix.Multiply(x, result) (SymmetricCSRMatrix.cs:293-324), nrhs columns. */
int mg_spmv_dev(mg_matrix *m, int32_t nrhs, const double *d_x, int32_t ldx, double *d_y, int32_t ldy,
                void *stream);
int mg_spmv(mg_matrix *m, const double *x, double *y);
/* ConjugateGradientGaussPreSolver.SolveWithInitialGuess (CGGP/ConjugateGradientGaussPreSolver.cs:78-166):
 * x0 may be NULL (x0 = 0); max_iterations = 0 means n; use_jacobi selects JacobiPreconditioner
 * (CGGP/JacobiPreconditioner.cs:12-56).  x receives the last iterate. */
int mg_pcg_solve(mg_matrix *m, const double *b, const double *x0, double relative_tolerance,
                 int32_t max_iterations, int32_t use_jacobi, double *x, mg_pcg_result *result);
int mg_pcg_solve_dev(mg_matrix *m, const double *d_b, const double *d_x0, double relative_tolerance,
                     int32_t max_iterations, int32_t use_jacobi, double *d_x, void *stream,
                     mg_pcg_result *result);

/* ---- general square CSR, ILU(0) and BiCGSTAB (BiCgStabSolver.cs, Ilu0Preconditioner.cs, CSRMatrix.cs) ---- */
/* CSRMatrix(int, int, IEnumerable<MatrixEntry>) (CSRMatrix.cs:34-80): duplicates of one (row, col)
 * summed in input order, |sum| <= 1e-15 dropped, columns ascending; indices checked as in
 * CSRMatrixBuilder.AddToElement (CSRMatrixBuilder.cs:74-81).  Two-call protocol as mg_csr_from_coo. */
int mg_gcsr_from_coo(int32_t rows, int32_t cols, int64_t count, const int32_t *row_idx, const int32_t *col_idx,
                     const double *vals, int32_t *nnz_out, int32_t *rowptr, int32_t *colidx, double *values);
/* CSRMatrix.LineCross (CSRMatrix.cs:110-164) for each of `count` equations in order; values and rhs in place. */
int mg_gcsr_line_cross(int32_t rows, int32_t cols, const int32_t *rowptr, const int32_t *colidx, double *values,
                       double *rhs, int32_t count, const int32_t *indices, const double *prescribed);
typedef struct mg_gmatrix mg_gmatrix;
/* device copy of a CSRMatrix (rows x cols, ascending columns per row) */
int mg_gmatrix_create(int32_t rows, int32_t cols, const int32_t *rowptr, const int32_t *colidx,
                      const double *values, mg_gmatrix **out);
void mg_gmatrix_release(mg_gmatrix *m);
/* CSRMatrix.Multiply (CSRMatrix.cs:230-248): y = A x, per-row sums in position order */
int mg_gspmv(mg_gmatrix *m, const double *x, double *y);
int mg_gspmv_dev(mg_gmatrix *m, const double *d_x, double *d_y, void *stream);
typedef struct mg_ilu0 mg_ilu0;
/* Ilu0Preconditioner(CSRMatrix, pivotTolerance) (BiCGSTAB/Ilu0Preconditioner.cs:14-31,86-166):
 * ILU(0) in the matrix's own pattern, level-scheduled on the GPU with the reference's order of
 * operations.  MG_E_ARG for a non-square matrix, MG_E_RANGE for a bad tolerance, MG_E_NOT_PD for a
 * missing diagonal or a zero/invalid pivot (the message names the first such row). */
int mg_ilu0_create(mg_gmatrix *m, double pivot_tolerance, mg_ilu0 **out);
void mg_ilu0_release(mg_ilu0 *p);
/* Ilu0Preconditioner.Apply (:44-84): x = (LU)^-1 b */
int mg_ilu0_apply(mg_ilu0 *p, const double *b, double *x);
int mg_ilu0_apply_dev(mg_ilu0 *p, const double *d_b, double *d_x, void *stream);
/* BiCgStabSolver.SolveWithInitialGuess (BiCGSTAB/BiCgStabSolver.cs:37-148): x0 may be NULL; pre may be
 * NULL (then use_preconditioner builds ILU(0) with pivot tolerance = breakdown_tolerance once the initial
 * residual has not passed); max_iterations = 0 means n.  Breakdowns return MG_E_NOT_PD with the
 * reference's message; x receives the last iterate. */
int mg_bicgstab_solve(mg_gmatrix *m, mg_ilu0 *pre, int32_t use_preconditioner, const double *b, const double *x0,
                      double relative_tolerance, int32_t max_iterations, double breakdown_tolerance, double *x,
                      mg_pcg_result *result);
int mg_bicgstab_solve_dev(mg_gmatrix *m, mg_ilu0 *pre, int32_t use_preconditioner, const double *d_b,
                          const double *d_x0, double relative_tolerance, int32_t max_iterations,
                          double breakdown_tolerance, double *d_x, void *stream, mg_pcg_result *result);

/* ---- batched independent systems ----
 * Many SymmetricCSRMatrix systems factored and solved together (the symbolic phase of each is
 * the reference's, UtduSymbolicFactorization.Analyze (UtduSymbolicFactorization.cs:240-295),
 * run over host threads).  The systems are merged into one block-diagonal plan: all fronts of
 * one tree level, from every system, go through the same launches; pivot scale and pivot rule
 * are per system (UtduNumericFactorization.cs:411-418,559-572).  Values and right-hand sides
 * are concatenated in system order (offsets from mg_batch_layout); a system with fewer
 * right-hand sides than nrhs passes zero columns.  Trailing zero columns cost no solve work:
 * each system's solves stop at its last nonzero column, and its solution columns past it are
 * set to zero. */
typedef struct mg_batch mg_batch;
int mg_batch_analyze(int32_t count, const int32_t *sizes, const int32_t *const *rowptrs,
                     const int32_t *const *colidxs, const mg_options *opts, mg_batch **out);
/* row_begin and nnz_begin (count+1 each, may be NULL); info of the merged factorization (may be NULL) */
int mg_batch_layout(const mg_batch *batch, int32_t *row_begin, int32_t *nnz_begin, mg_symbolic_info *info);
int mg_batch_factorize_dev(mg_batch *batch, const double *d_values, const mg_options *opts, void *stream);
/* every system and column refined until its own ||b-Ax||/||b|| <= tol (report: worst over all) */
int mg_batch_solve_dev(mg_batch *batch, int32_t nrhs, const double *d_b, int32_t ldb, double *d_x, int32_t ldx,
                       double tol, int32_t max_refine, void *stream, mg_solve_report *report);
/* host arrays per system: values[t] (nnz of system t), b[t] and x[t] column-major n_t x nrhs[t] */
int mg_batch_factorize(mg_batch *batch, const double *const *values, const mg_options *opts);
int mg_batch_solve(mg_batch *batch, const int32_t *nrhs, const double *const *b, double *const *x, double tol,
                   int32_t max_refine, mg_solve_report *report);
int mg_batch_numeric_info(const mg_batch *batch, mg_numeric_info *info);
/* new: Schur tile autotuner.  The trailing (outer) Schur updates of a front run with 64x64
 * or 128x128 MFMA output tiles per front-size bucket (untuned: 128 for fronts of 1024 rows
 * or more); the narrow updates inside a front's column blocks always use 64x64.  For each
 * bucket with outer updates it times the whole factorization of d_values with 64x64 and
 * with 128x128 tiles (best of three runs each, the other buckets at their current choice)
 * and keeps 128 only when it is at least 2% faster; the choice is stored with the symbolic
 * factorization for the handle's precision and used by every later factorization (a
 * captured graph is re-captured).  The handle is left factorized from d_values.  tiles
 * (MG_SCHUR_BUCKETS entries, may be NULL) receives 64 or 128 per bucket; ms (may be NULL)
 * the tuned factorization time. */
int mg_autotune_dev(mg_numeric *num, const double *d_values, const mg_options *opts, void *stream,
                    int32_t *tiles, double *ms);
int mg_batch_autotune_dev(mg_batch *batch, const double *d_values, const mg_options *opts, void *stream,
                          int32_t *tiles, double *ms);
void mg_batch_release(mg_batch *batch);

/* ---- learned preconditioner ----
 * No counterpart in the reference: a message-passing network on the matrix graph (three
 * layers of width 32) predicts a sparse lower-triangular factor G (columns within two hops),
 * and z = G'G r corrects JacobiPreconditioner (CGGP/JacobiPreconditioner.cs:12-56; zero
 * readouts give Jacobi) inside a flexible PCG.  It is trained on the GPU from (residual, exact
 * error) pairs that the direct solver produces, with a Kaporin condition-number term of
 * weight kappa; inference runs the network once per matrix with BF16 weights and FP32
 * accumulation on MFMA.  The solve falls back to plain Jacobi-PCG when it stalls. */
typedef struct mg_mpnn mg_mpnn;
int32_t mg_mpnn_param_count(void);
int mg_mpnn_create(uint64_t seed, mg_mpnn **out);
void mg_mpnn_release(mg_mpnn *net);
int mg_mpnn_get_params(const mg_mpnn *net, float *params);   /* FP32 master copy, param_count values */
int mg_mpnn_set_params(mg_mpnn *net, const float *params);   /* also refreshes the BF16 copy */
/* z = G'G r with G from the FP32 network (MG_FP32) or the BF16 MFMA network (MG_BF16) */
int mg_mpnn_apply_dev(mg_mpnn *net, mg_matrix *m, const double *d_r, double *d_z, int32_t precision,
                      void *stream);
/* mean A-norm loss over ns pairs (columns of d_R = residuals and d_E = exact errors, leading
 * dimension ld) plus kappa times the log Kaporin number of G A G', and its gradient (host array
 * of param_count floats); no update */
int mg_mpnn_loss_grad_dev(mg_mpnn *net, mg_matrix *m, int32_t ns, const double *d_R, const double *d_E,
                          int64_t ld, double kappa, void *stream, double *loss, float *grad);
/* Adam training, `epochs` passes over the pairs in seeded shuffled mini-batches; epoch_losses
 * (may be NULL) receives the mean loss of each epoch */
int mg_mpnn_train_dev(mg_mpnn *net, mg_matrix *m, int32_t ns, const double *d_R, const double *d_E, int64_t ld,
                      int32_t epochs, int32_t batch, float learning_rate, double kappa, uint64_t seed, void *stream,
                      double *epoch_losses);
/* SolveWithInitialGuess (CGGP/ConjugateGradientGaussPreSolver.cs:78-166) with the learned
 * preconditioner: same stopping rule and result; *fell_back = 1 when it switched to Jacobi
 * (r'z <= 0, or no new minimum of ||r|| for stall_window iterations; 0 selects 50). */
int mg_pcg_solve_learned_dev(mg_matrix *m, mg_mpnn *net, const double *d_b, const double *d_x0,
                             double relative_tolerance, int32_t max_iterations, int32_t stall_window, double *d_x,
                             void *stream, mg_pcg_result *result, int32_t *fell_back);
int mg_pcg_solve_learned(mg_matrix *m, mg_mpnn *net, const double *b, const double *x0, double relative_tolerance,
                         int32_t max_iterations, int32_t stall_window, double *x, mg_pcg_result *result,
                         int32_t *fell_back);

#ifdef __cplusplus
}
#endif
#endif
