/*
 * oracle.h -- CPU restatement of MAGLA's UCHOL / CGGP path, used ONLY as a
 * test checker and as the CPU timing baseline.  Nothing in the product
 * library (magla_amd/) links, loads or calls this code.
 *
 * Every function follows a reference file (paths relative to the MAGLA
 * repository root, CAESolvers/...); the exact lines are cited at each
 * definition in oracle.c.  All arithmetic is IEEE binary64 in the
 * reference's operation order; build with -ffp-contract=off (RyuJIT never
 * fuses a multiply into an add).
 */
#ifndef MAGLA_ORACLE_H
#define MAGLA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes: the same values as the product ABI (include/magla.h) */
#define OR_OK 0
#define OR_E_NULL 1
#define OR_E_ARG 2
#define OR_E_RANGE 3
#define OR_E_NOT_PD 4
#define OR_E_TOO_DENSE 5
#define OR_E_INTERNAL 7
#define OR_E_NOMEM 8

/* ---- symmetric CSR (upper triangle incl. diagonal) ---- */
typedef struct {
    int n;
    int nnz;
    int *rowptr;   /* n+1 */
    int *colidx;   /* nnz, ascending inside a row */
    double *values;/* nnz */
    int *diag;     /* n, position of A[i,i] or -1 */
} or_csr;

/* COO -> CSR: SymmetricCSRMatrix(int, IEnumerable<MatrixEntry>) */
int or_csr_from_coo(int n, int64_t count, const int *rows, const int *cols,
                    const double *vals, or_csr **out);
/* wrap existing arrays (copied) */
int or_csr_create(int n, const int *rowptr, const int *colidx, const double *values,
                  or_csr **out);
void or_csr_free(or_csr *a);
/* y = A x, SymmetricCSRMatrix.Multiply(double[], double[]) */
void or_spmv(const or_csr *a, const double *x, double *y);
/* SymmetricCSRMatrix.LineCross */
int or_linecross(or_csr *a, double *rhs, double prescribed, int index);

/* ---- symbolic phase ---- */
typedef struct {
    int ordering;              /* 0 natural, 1 AMD */
    int amalgamation_limit;    /* 64 */
    int panel_width;           /* 32 (clamped 4..128) */
    int allow_regularization;  /* 1 */
    double pivot_tolerance;    /* 1e-13 */
    double regularization;     /* 1e-10 */
    double time_budget_s;      /* 0 = factor everything; else stop after the supernode that
                                  crosses this wall time (bounded CPU-baseline sample) */
} or_options;

void or_options_default(or_options *o);

typedef struct {
    int n;
    int matrix_nnz;
    int *amd;          /* fill-reducing order before postorder (n) */
    int *perm;         /* perm[k] = original index eliminated k-th */
    int *iperm;
    int *etree;        /* final elimination tree (permuted numbering) */
    int *colcount;     /* final column counts */
    /* permuted pattern (PermutedSymmetricPattern) */
    int *pptr;         /* n+1 */
    int *prows;        /* full structure, no diagonal, ascending */
    int *lower_start;  /* n */
    int *lvptr;        /* n+1 */
    int *lvidx;        /* lower value indices */
    int *dvidx;        /* diagonal value indices (-1 if absent) */
    /* supernodes */
    int nsuper;
    int *first;        /* nsuper+1 */
    int *super_of;     /* n */
    int *sparent;      /* nsuper */
    int *child_ptr;    /* nsuper+1 */
    int *children;
    int *pat_ptr;      /* nsuper+1 */
    int *pat_rows;
    int64_t *block_off;/* nsuper+1 : packed factor block offsets */
    int64_t factor_entries;
    int64_t strict_nonzeros;
    int64_t factor_ops;
    int max_front;
    int max_width;
    int64_t peak_contribution;
} or_symbolic;

int or_amd(int n, const int *rowptr, const int *colidx, int *perm_out);
int or_analyze(const or_csr *a, const or_options *o, or_symbolic **out);
void or_symbolic_free(or_symbolic *s);

/* ---- numeric phase ---- */
typedef struct {
    const or_symbolic *sym;
    double *factor;    /* packed blocks, block_off[s] .. */
    double *diag;      /* n, elimination order */
    int regularized;
    int negative;
    double smallest;
    double largest;
    int supernodes_done;       /* < nsuper only when time_budget_s cut the run short */
    int64_t ops_done;          /* FactorOperationCount of the supernodes factored */
} or_numeric;

int or_factorize(const or_symbolic *s, const or_csr *a, const or_options *o,
                 or_numeric **out);
void or_numeric_free(or_numeric *f);
/* UtduNumericFactorization.Solve: x = A^-1 b (b unchanged) */
int or_solve(const or_numeric *f, const double *b, double *x);
/* The reference's threaded factorization (tree schedule over `threads` workers, parallel
 * column chunks of large trailing updates, blocked trailing update): the same factor as
 * or_factorize, bit for bit, for any thread count (CPU timing baseline). */
int or_factorize_mt(const or_symbolic *s, const or_csr *a, const or_options *o, int threads,
                    or_numeric **out);
/* nrhs right-hand sides (column-major, leading dimension n) solved over `threads` threads */
int or_solve_mt(const or_numeric *f, int nrhs, const double *b, double *x, int threads);

/* ---- CGGP ---- */
typedef struct {
    int iterations;
    int converged;
    double residual_norm;
    double relative_residual;
} or_pcg_result;

/* use_jacobi: 1 = JacobiPreconditioner, 0 = identity.  x0 may be NULL. */
int or_pcg(const or_csr *a, const double *b, const double *x0, double tol,
           int max_iterations, int use_jacobi, double *x, or_pcg_result *res);

double or_norm2(int n, const double *v);

/* ---- learned preconditioner network (checker for the HIP training/inference kernels) ----
 * The reference has no learned component: these functions restate the preconditioner that this
 * project defines, in binary64, so that the GPU forward pass and gradients can be checked.
 * A message-passing network (three layers of width 32) on the matrix graph predicts a sparse
 * lower-triangular factor G on the pattern of tril(A + A^2); the preconditioner is z = G'G r,
 * a correction on top of Jacobi (G = D^{-1/2} when the readouts are zero).
 *   edge weights   e1_ij = a_ij / sqrt(a_ii a_jj) (j != i);  e2_ij = sum_k e1_ik e1_kj (j != i)
 *   node features  x_i = [1, c_i, s_i], c_i = log(a_ii / mean a), s_i = sum_j |e1_ij|
 *   h0_i = W_in x_i ;  layer l: m_i = sum_j e1_ij h_j, h'_i = act(W1_l h_i + W2_l m_i)
 *                      (act = ReLU for l < 2, identity for l = 2)
 *   t_i  = w_t . h3_i                                  G_ii = a_ii^{-1/2} exp(t_i)
 *   p_ij = u . (h3_i * h3_j) + v . h3_i + w . h3_j     with (u, v, w) = (u1, v1, w1) or (u2, v2, w2)
 *   G_ij = a_jj^{-1/2} (e1_ij p1_ij + e2_ij p2_ij)     (j < i, j within two hops of i)
 * Parameter layout, OR_MP_P values: W_in[32][3] | for l = 0..2: W1_l[32][32], W2_l[32][32]
 * (row f, column g) | w_t | u1 | v1 | w1 | u2 | v2 | w2 (32 each).
 * Loss: the mean over (r, e = A^{-1} r) pairs of the squared A-norm error of z relative to e,
 *   L_pair = (z - e)' A (z - e) / (e' A e) = (z'Az - 2 z'r) / (r'e) + 1,
 * plus kappa times the log of Kaporin's condition number of G A G' (up to the constant log det A),
 *   L_K = log(tr(G A G') / n) - (2/n) sum_i log G_ii,
 * which keeps the small eigenvalues of G A G' away from zero (the pair term alone does not). */
#define OR_MP_W 32
#define OR_MP_P (3 * OR_MP_W + 3 * 2 * OR_MP_W * OR_MP_W + 7 * OR_MP_W)
int or_mpnn_apply(const or_csr *a, const double *params, const double *r, double *z);
/* mean loss over nsamples pairs (columns of R and E, length n each) and its gradient */
int or_mpnn_loss_grad(const or_csr *a, const double *params, int nsamples, const double *R, const double *E,
                      double kappa, double *loss, double *grad);

/* ---- general CSR, ILU(0), BiCGSTAB (CSRMatrix.cs, BiCGSTAB/BiCgStabSolver.cs,
 *      BiCGSTAB/Ilu0Preconditioner.cs) ---- */
typedef struct {
    int rows, cols, nnz;
    int *rowptr, *colidx;
    double *values;
    int *diag;     /* rows, position of A[i,i] or -1 */
} or_gcsr;
/* CSRMatrix(int, int, IEnumerable<MatrixEntry>): duplicates summed in input order, |sum| <= 1e-15
 * dropped, columns ascending */
int or_gcsr_from_coo(int rows, int cols, int64_t count, const int *ri, const int *cj, const double *vals,
                     or_gcsr **out);
int or_gcsr_create(int rows, int cols, const int *rowptr, const int *colidx, const double *values, or_gcsr **out);
void or_gcsr_free(or_gcsr *a);
/* CSRMatrix.Multiply */
void or_gspmv(const or_gcsr *a, const double *x, double *y);
/* CSRMatrix.LineCross */
int or_gcsr_linecross(or_gcsr *a, double *rhs, double prescribed, int index);
typedef struct {
    const or_gcsr *a;
    double pivot_tolerance;
    double *factors; /* nnz, L (unit, strictly lower) and U in A's pattern */
    int *diag;
} or_ilu0;
/* Ilu0Preconditioner(CSRMatrix, pivotTolerance); *bad_row receives the failing row (or -1) */
int or_ilu0_create(const or_gcsr *a, double pivot_tolerance, or_ilu0 **out, int *bad_row);
void or_ilu0_apply(const or_ilu0 *p, const double *b, double *x);
void or_ilu0_free(or_ilu0 *p);
/* BiCgStabSolver.SolveWithInitialGuess; pre may be NULL (use_pre builds ILU(0) with pivot tolerance
 * btol); a breakdown returns OR_E_NOT_PD with *fail_code 1 rho, 2 omega (before beta),
 * 3 (r0, A M^-1 p), 4 (A M^-1 s, A M^-1 s), 5 omega, and *fail_it its iteration; an ILU(0)
 * failure returns OR_E_NOT_PD with *fail_code 6 (missing diagonal) or 7 (pivot). */
int or_bicgstab(const or_gcsr *a, const or_ilu0 *pre, int use_pre, const double *b, const double *x0, double tol,
                int max_iterations, double btol, double *x, or_pcg_result *res, int *fail_code, int *fail_it);

#ifdef __cplusplus
}
#endif
#endif
