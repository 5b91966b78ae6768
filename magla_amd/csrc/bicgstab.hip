// bicgstab.hip -- BiCgStabSolver with Ilu0Preconditioner for general square CSR matrices
// (BiCGSTAB/BiCgStabSolver.cs, BiCGSTAB/Ilu0Preconditioner.cs, CSRMatrix.cs) on the GPU.
//
// ILU(0) and its two triangular solves are level-scheduled: row i of L belongs to level
// 1 + max(level of the rows of its strictly lower columns), and the rows of one level are
// independent.  One workgroup walks the levels with a barrier between them; each row is
// done by one thread in the reference's own order of operations (positions ascending, no
// fused multiply-add), so factors and solves equal the reference's sequential loops bit for
// bit.  The general SpMV is one thread per row in position order for the same reason.  The
// BiCGSTAB iteration keeps its scalars on the device; dot products use fixed-order block
// reductions.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "device.h"

namespace mg {

namespace {

constexpr int kLevelThreads = 1024;

// reference arithmetic: products and sums rounded separately
#define MG_NO_CONTRACT _Pragma("clang fp contract(off)")

__global__ __launch_bounds__(kLevelThreads) void k_ilu0_factor(const int *__restrict__ rp, const int *__restrict__ ci,
                                                               const int *__restrict__ diag, double *__restrict__ f,
                                                               const int *__restrict__ lptr,
                                                               const int *__restrict__ lrows, int nlev, double tol,
                                                               int *__restrict__ bad) {
    MG_NO_CONTRACT
    for (int L = 0; L < nlev; ++L) {
        for (int q = lptr[L] + (int)threadIdx.x; q < lptr[L + 1]; q += kLevelThreads) {
            const int row = lrows[q], end = rp[row + 1], dr = diag[row];
            // Factorize (Ilu0Preconditioner.cs:111-137): for each lower position in order
            for (int p = rp[row]; p < dr; ++p) {
                const int col = ci[p];
                const double mult = f[p] / f[diag[col]];
                f[p] = mult;
                // (row, ci[u]) for the upper positions u of row col; the target lies after p
                // and the columns ascend, so a forward walk finds what FindPosition finds
                int t = p + 1;
                for (int u = diag[col] + 1; u < rp[col + 1]; ++u) {
                    const int j = ci[u];
                    while (t < end && ci[t] < j) ++t;
                    if (t < end && ci[t] == j) f[t] = f[t] - mult * f[u];
                }
            }
            const double d = f[dr];
            if (!isfinite(d) || fabs(d) < tol) atomicMin(bad, row);  // EnsurePivot (:162-166)
        }
        __syncthreads();
    }
}

// w = L^{-1} r (Ilu0Preconditioner.cs:67-74)
__global__ __launch_bounds__(kLevelThreads) void k_ilu0_lower(const int *__restrict__ rp, const int *__restrict__ ci,
                                                              const int *__restrict__ diag,
                                                              const double *__restrict__ f,
                                                              const int *__restrict__ lptr,
                                                              const int *__restrict__ lrows, int nlev,
                                                              const double *__restrict__ r, double *__restrict__ w,
                                                              const int *__restrict__ done) {
    MG_NO_CONTRACT
    if (done && *done) return;
    for (int L = 0; L < nlev; ++L) {
        for (int q = lptr[L] + (int)threadIdx.x; q < lptr[L + 1]; q += kLevelThreads) {
            const int row = lrows[q];
            double sum = r[row];
            for (int p = rp[row]; p < diag[row]; ++p) sum -= f[p] * w[ci[p]];
            w[row] = sum;
        }
        __syncthreads();
    }
}

// x = U^{-1} w (Ilu0Preconditioner.cs:76-83)
__global__ __launch_bounds__(kLevelThreads) void k_ilu0_upper(const int *__restrict__ rp, const int *__restrict__ ci,
                                                              const int *__restrict__ diag,
                                                              const double *__restrict__ f,
                                                              const int *__restrict__ uptr,
                                                              const int *__restrict__ urows, int nlev,
                                                              const double *__restrict__ w, double *__restrict__ x,
                                                              const int *__restrict__ done) {
    MG_NO_CONTRACT
    if (done && *done) return;
    for (int L = 0; L < nlev; ++L) {
        for (int q = uptr[L] + (int)threadIdx.x; q < uptr[L + 1]; q += kLevelThreads) {
            const int row = urows[q], dr = diag[row];
            double sum = w[row];
            for (int p = dr + 1; p < rp[row + 1]; ++p) sum -= f[p] * x[ci[p]];
            x[row] = sum / f[dr];
        }
        __syncthreads();
    }
}

// y = A x, one thread per row in position order (BiCgStabSolver.cs:167-180, CSRMatrix.cs:230-248)
__global__ __launch_bounds__(256) void k_gspmv(const int *__restrict__ rp, const int *__restrict__ ci,
                                               const double *__restrict__ v, int rows, const double *__restrict__ x,
                                               double *__restrict__ y, const int *__restrict__ done) {
    MG_NO_CONTRACT
    if (done && *done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    double sum = 0.0;
    for (int p = rp[i]; p < rp[i + 1]; ++p) sum += v[p] * x[ci[p]];
    y[i] = sum;
}

// ---- BiCGSTAB (BiCgStabSolver.cs:37-148) ----

struct BicgScalars {
    double rho, rho_prev, alpha, omega, beta, rnorm, snorm, thr, bnorm;
    int done, iters, fail, fail_it, early, pad;
};

constexpr int kBgBS = 256;
constexpr int kBgParts = 512;  // fixed grid of the reductions

__device__ inline double bg_block_sum(double v, double *sm) {
    sm[threadIdx.x] = v;
    __syncthreads();
    for (int h = kBgBS / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sm[threadIdx.x] += sm[threadIdx.x + h];
        __syncthreads();
    }
    return sm[0];
}

__device__ inline double bg_finish(const double *part, double *sm) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < kBgParts; b += kBgBS) acc += part[b];
    return bg_block_sum(acc, sm);
}

__device__ inline bool bg_breakdown(double v, double tol) { return !isfinite(v) || fabs(v) < tol; }

__device__ inline void bg_fail(BicgScalars *S, int code, int it) {
    S->fail = code;
    S->fail_it = it;
    S->done = 1;
}

__global__ __launch_bounds__(kBgBS) void k_bg_dot(const double *__restrict__ a, const double *__restrict__ c, int n,
                                                  double *__restrict__ part, const BicgScalars *__restrict__ S) {
    __shared__ double sm[kBgBS];
    if (S && S->done) return;
    double acc = 0.0;
    for (int i = blockIdx.x * kBgBS + threadIdx.x; i < n; i += gridDim.x * kBgBS) acc += a[i] * c[i];
    const double t = bg_block_sum(acc, sm);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// r = b - A x (:69-73)
__global__ void k_bg_residual0(const double *__restrict__ b, const double *__restrict__ ax, int n,
                               double *__restrict__ r, double *__restrict__ rhat) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        r[i] = b[i] - ax[i];
        rhat[i] = r[i];
    }
}

// residualScale, threshold and the initial test (:54-79)
__global__ void k_bg_init(const double *__restrict__ prr, const double *__restrict__ pbb, double tol,
                          BicgScalars *__restrict__ S) {
    __shared__ double sm[kBgBS];
    const double rr = bg_finish(prr, sm);
    __syncthreads();
    const double bb = bg_finish(pbb, sm);
    if (threadIdx.x == 0) {
        const double bn = sqrt(bb);
        S->bnorm = bn > 0.0 ? bn : 1.0;
        S->thr = tol * S->bnorm;
        S->rnorm = sqrt(rr);
        S->done = S->rnorm <= S->thr ? 1 : 0;
        S->iters = 0;
        S->fail = 0;
        S->early = 0;
        S->rho_prev = 1.0;
        S->alpha = 1.0;
        S->omega = 1.0;
    }
}

// rho, the breakdown tests and beta (:92-102)
__global__ void k_bg_rho(const double *__restrict__ part, int it, double btol, BicgScalars *__restrict__ S) {
    __shared__ double sm[kBgBS];
    if (S->done) return;
    const double rho = bg_finish(part, sm);
    if (threadIdx.x == 0) {
        S->rho = rho;
        if (bg_breakdown(rho, btol)) {
            bg_fail(S, 1, it);
        } else if (it > 1) {
            if (bg_breakdown(S->omega, btol)) bg_fail(S, 2, it);
            else S->beta = rho / S->rho_prev * (S->alpha / S->omega);
        }
    }
}

// p = r (first iteration) or r + beta (p - omega v) (:95-105)
__global__ void k_bg_direction(const double *__restrict__ r, double *__restrict__ p, const double *__restrict__ v,
                               int n, int it, const BicgScalars *__restrict__ S) {
    MG_NO_CONTRACT
    if (S->done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    p[i] = it == 1 ? r[i] : r[i] + S->beta * (p[i] - S->omega * v[i]);
}

__global__ void k_bg_copy(const double *__restrict__ a, double *__restrict__ c, int n,
                          const BicgScalars *__restrict__ S) {
    if (S->done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c[i] = a[i];
}

// alpha = rho / (r0, A M^-1 p) (:110-112)
__global__ void k_bg_alpha(const double *__restrict__ part, int it, double btol, BicgScalars *__restrict__ S) {
    __shared__ double sm[kBgBS];
    if (S->done) return;
    const double den = bg_finish(part, sm);
    if (threadIdx.x == 0) {
        if (bg_breakdown(den, btol)) bg_fail(S, 3, it);
        else S->alpha = S->rho / den;
    }
}

// s = r - alpha v and its squared norm (:114-117)
__global__ __launch_bounds__(kBgBS) void k_bg_s(const double *__restrict__ r, const double *__restrict__ v,
                                                double *__restrict__ s, int n, double *__restrict__ part,
                                                const BicgScalars *__restrict__ S) {
    MG_NO_CONTRACT
    __shared__ double sm[kBgBS];
    if (S->done) return;
    const double alpha = S->alpha;
    double acc = 0.0;
    for (int i = blockIdx.x * kBgBS + threadIdx.x; i < n; i += gridDim.x * kBgBS) {
        const double si = r[i] - alpha * v[i];
        s[i] = si;
        acc += si * si;
    }
    const double t = bg_block_sum(acc, sm);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// ||s|| <= threshold: x += alpha M^-1 p and stop (:118-124)
__global__ void k_bg_snorm(const double *__restrict__ part, BicgScalars *__restrict__ S) {
    __shared__ double sm[kBgBS];
    if (S->done) return;
    const double ss = bg_finish(part, sm);
    if (threadIdx.x == 0) {
        S->snorm = sqrt(ss);
        S->early = S->snorm <= S->thr ? 1 : 0;
    }
}

__global__ void k_bg_xearly(double *__restrict__ x, const double *__restrict__ ph, int n,
                            const BicgScalars *__restrict__ S) {
    MG_NO_CONTRACT
    if (S->done || !S->early) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += S->alpha * ph[i];
}

__global__ void k_bg_early_done(int it, BicgScalars *__restrict__ S) {
    if (threadIdx.x != 0 || S->done || !S->early) return;
    S->done = 1;
    S->iters = it;
    S->rnorm = S->snorm;
}

// omega = (t, s) / (t, t) with its two breakdown tests (:129-132)
__global__ void k_bg_omega(const double *__restrict__ ptt, const double *__restrict__ pts, int it, double btol,
                           BicgScalars *__restrict__ S) {
    __shared__ double sm[kBgBS];
    if (S->done) return;
    const double tt = bg_finish(ptt, sm);
    __syncthreads();
    const double ts = bg_finish(pts, sm);
    if (threadIdx.x == 0) {
        if (bg_breakdown(tt, btol)) {
            bg_fail(S, 4, it);
        } else {
            const double om = ts / tt;
            S->omega = om;
            if (bg_breakdown(om, btol)) bg_fail(S, 5, it);
        }
    }
}

// x += alpha M^-1 p + omega M^-1 s; r = s - omega t; ||r||^2 (:134-140)
__global__ __launch_bounds__(kBgBS) void k_bg_update(double *__restrict__ x, double *__restrict__ r,
                                                     const double *__restrict__ ph, const double *__restrict__ sh,
                                                     const double *__restrict__ s, const double *__restrict__ t,
                                                     int n, double *__restrict__ part,
                                                     const BicgScalars *__restrict__ S) {
    MG_NO_CONTRACT
    __shared__ double sm[kBgBS];
    if (S->done) return;
    const double alpha = S->alpha, omega = S->omega;
    double acc = 0.0;
    for (int i = blockIdx.x * kBgBS + threadIdx.x; i < n; i += gridDim.x * kBgBS) {
        x[i] += alpha * ph[i] + omega * sh[i];
        const double ri = s[i] - omega * t[i];
        r[i] = ri;
        acc += ri * ri;
    }
    const double tv = bg_block_sum(acc, sm);
    if (threadIdx.x == 0) part[blockIdx.x] = tv;
}

// ||r|| <= threshold ends the iteration (:140-144)
__global__ void k_bg_rnorm(const double *__restrict__ part, int it, BicgScalars *__restrict__ S) {
    __shared__ double sm[kBgBS];
    if (S->done) return;
    const double rr = bg_finish(part, sm);
    if (threadIdx.x == 0) {
        S->rnorm = sqrt(rr);
        S->iters = it;
        if (S->rnorm <= S->thr) S->done = 1;
        S->rho_prev = S->rho;
    }
}

inline unsigned grid_of(int n, int bs) { return (unsigned)std::max(1, (n + bs - 1) / bs); }

}  // namespace

// ---- host side ----

void gmatrix_build(GMatrix &G, int rows, int cols, const int *rowptr, const int *colidx, const double *values) {
    if (rows < 0 || cols < 0) throw Error(MG_E_ARG, "Matrix dimensions cannot be negative.");
    if (!rowptr) throw Error(MG_E_NULL, "rowptr is null.");
    if (rowptr[0] != 0) throw Error(MG_E_ARG, "rowptr[0] must be 0.");
    for (int r = 0; r < rows; ++r) {
        if (rowptr[r + 1] < rowptr[r]) throw Error(MG_E_ARG, "rowptr must be non-decreasing.");
        for (int p = rowptr[r]; p < rowptr[r + 1]; ++p) {
            if (colidx[p] < 0 || colidx[p] >= cols) throw Error(MG_E_RANGE, "Column index is out of range.");
            if (p > rowptr[r] && colidx[p] <= colidx[p - 1])
                throw Error(MG_E_ARG, "Column indices must ascend within a row.");
        }
    }
    G.rows = rows;
    G.cols = cols;
    G.nnz = rowptr[rows];
    G.h_rowptr.assign(rowptr, rowptr + rows + 1);
    G.h_colidx.assign(colidx, colidx + G.nnz);
    G.rowptr.upload(G.h_rowptr);
    G.colidx.upload(G.h_colidx);
    G.vals.alloc(std::max(G.nnz, 1));
    if (G.nnz) MG_HIP(hipMemcpy(G.vals.p, values, sizeof(double) * G.nnz, hipMemcpyHostToDevice));
}

void gspmv_device(const GMatrix &G, const double *d_x, double *d_y, hipStream_t st) {
    if (G.rows == 0) return;
    hipLaunchKernelGGL(k_gspmv, dim3(grid_of(G.rows, 256)), dim3(256), 0, st, G.rowptr.p, G.colidx.p, G.vals.p,
                       G.rows, d_x, d_y, (const int *)nullptr);
    MG_HIP(hipGetLastError());
}

// level of each row: lower = 1 + max over strictly lower columns, upper likewise from the end
static void build_levels(int n, const std::vector<int> &rp, const std::vector<int> &ci, const std::vector<int> &dg,
                         bool lower, std::vector<int> &ptr, std::vector<int> &rows) {
    std::vector<int> lev(n, 0);
    int nlev = 0;
    if (lower) {
        for (int i = 0; i < n; ++i) {
            int l = 0;
            for (int p = rp[i]; p < dg[i]; ++p) l = std::max(l, lev[ci[p]] + 1);
            lev[i] = l;
            nlev = std::max(nlev, l + 1);
        }
    } else {
        for (int i = n - 1; i >= 0; --i) {
            int l = 0;
            for (int p = dg[i] + 1; p < rp[i + 1]; ++p) l = std::max(l, lev[ci[p]] + 1);
            lev[i] = l;
            nlev = std::max(nlev, l + 1);
        }
    }
    ptr.assign(nlev + 1, 0);
    for (int i = 0; i < n; ++i) ++ptr[lev[i] + 1];
    for (int l = 0; l < nlev; ++l) ptr[l + 1] += ptr[l];
    rows.assign(n, 0);
    std::vector<int> at(ptr.begin(), ptr.end() - 1);
    for (int i = 0; i < n; ++i) rows[at[lev[i]]++] = i;  // ascending within a level
}

void ilu0_build(Ilu0 &P, const GMatrix &A, double pivot_tol, hipStream_t st) {
    // Ilu0Preconditioner(CSRMatrix, double) (Ilu0Preconditioner.cs:14-31)
    if (A.rows != A.cols) throw Error(MG_E_ARG, "ILU(0) can be applied only to a square matrix.");
    if (!std::isfinite(pivot_tol) || pivot_tol <= 0.0) throw Error(MG_E_RANGE, "pivotTolerance is out of range.");
    const int n = A.rows;
    P.A = &A;
    P.n = n;
    P.pivot_tol = pivot_tol;
    std::vector<int> dg(n, -1);
    for (int i = 0; i < n; ++i) {  // FindDiagonalIndices (:86-109)
        for (int p = A.h_rowptr[i]; p < A.h_rowptr[i + 1]; ++p)
            if (A.h_colidx[p] == i) {
                dg[i] = p;
                break;
            }
        if (dg[i] < 0)
            throw Error(MG_E_NOT_PD,
                        "Row " + std::to_string(i) + " does not contain the diagonal element required for ILU(0).");
    }
    std::vector<int> lptr, lrows, uptr, urows;
    build_levels(n, A.h_rowptr, A.h_colidx, dg, true, lptr, lrows);
    build_levels(n, A.h_rowptr, A.h_colidx, dg, false, uptr, urows);
    P.nlev_l = (int)lptr.size() - 1;
    P.nlev_u = (int)uptr.size() - 1;
    P.diag.upload(dg);
    P.lptr.upload(lptr);
    P.lrows.upload(lrows);
    P.uptr.upload(uptr);
    P.urows.upload(urows);
    P.f.alloc(std::max(A.nnz, 1));
    P.w.alloc(std::max(n, 1));
    P.bad.alloc(1);
    if (n == 0) return;
    if (A.nnz) MG_HIP(hipMemcpyAsync(P.f.p, A.vals.p, sizeof(double) * A.nnz, hipMemcpyDeviceToDevice, st));
    const int none = 0x7fffffff;
    MG_HIP(hipMemcpyAsync(P.bad.p, &none, sizeof(int), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_ilu0_factor, dim3(1), dim3(kLevelThreads), 0, st, A.rowptr.p, A.colidx.p, P.diag.p, P.f.p,
                       P.lptr.p, P.lrows.p, P.nlev_l, pivot_tol, P.bad.p);
    MG_HIP(hipGetLastError());
    int bad = none;
    MG_HIP(hipMemcpyAsync(&bad, P.bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
    MG_HIP(hipStreamSynchronize(st));
    if (bad != none)
        throw Error(MG_E_NOT_PD,
                    "U has a zero or invalid pivot in row " + std::to_string(bad) + " while building ILU(0).");
}

void ilu0_apply_device(const Ilu0 &P, const double *d_r, double *d_z, hipStream_t st, const int *done) {
    if (P.n == 0) return;
    const GMatrix &A = *P.A;
    hipLaunchKernelGGL(k_ilu0_lower, dim3(1), dim3(kLevelThreads), 0, st, A.rowptr.p, A.colidx.p, P.diag.p, P.f.p,
                       P.lptr.p, P.lrows.p, P.nlev_l, d_r, P.w.p, done);
    hipLaunchKernelGGL(k_ilu0_upper, dim3(1), dim3(kLevelThreads), 0, st, A.rowptr.p, A.colidx.p, P.diag.p, P.f.p,
                       P.uptr.p, P.urows.p, P.nlev_u, P.w.p, d_z, done);
    MG_HIP(hipGetLastError());
}

void bicgstab_device(const GMatrix &A, const Ilu0 *pre, int use_pre, const double *d_b, const double *d_x0,
                     double tol, int max_it, double btol, double *d_x, hipStream_t st, mg_pcg_result &res,
                     std::unique_ptr<Ilu0> &owned) {
    // validation order of SolveWithInitialGuess (BiCgStabSolver.cs:39-51)
    if (!std::isfinite(tol) || tol < 0.0) throw Error(MG_E_RANGE, "RelativeTolerance is out of range.");
    if (max_it < 0) throw Error(MG_E_RANGE, "MaxIterations: the maximum iteration count cannot be negative.");
    if (!std::isfinite(btol) || btol <= 0.0) throw Error(MG_E_RANGE, "BreakdownTolerance is out of range.");
    if (A.rows != A.cols) throw Error(MG_E_ARG, "BiCGStab can be applied only to a square matrix.");
    if (pre && pre->A != &A) throw Error(MG_E_ARG, "The preconditioner was built for another matrix.");
    const int n = A.rows;
    res = mg_pcg_result{0, 1, 0.0, 0.0};
    if (n == 0) return;
    if (d_x0) MG_HIP(hipMemcpyAsync(d_x, d_x0, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    else MG_HIP(hipMemsetAsync(d_x, 0, sizeof(double) * n, st));
    DevArray<double> r(n), rh(n), p(n), ph(n), v(n), s(n), sh(n), t(n), pa(kBgParts), pb(kBgParts);
    DevArray<BicgScalars> S(1);
    const dim3 gv(grid_of(n, kBgBS)), bv(kBgBS), gr(kBgParts);
    const int *done = reinterpret_cast<const int *>(reinterpret_cast<const char *>(S.p) + offsetof(BicgScalars, done));
    hipLaunchKernelGGL(k_gspmv, gv, bv, 0, st, A.rowptr.p, A.colidx.p, A.vals.p, n, d_x, v.p, (const int *)nullptr);
    hipLaunchKernelGGL(k_bg_residual0, gv, bv, 0, st, d_b, v.p, n, r.p, rh.p);
    hipLaunchKernelGGL(k_bg_dot, gr, bv, 0, st, r.p, r.p, n, pa.p, (const BicgScalars *)nullptr);
    hipLaunchKernelGGL(k_bg_dot, gr, bv, 0, st, d_b, d_b, n, pb.p, (const BicgScalars *)nullptr);
    hipLaunchKernelGGL(k_bg_init, dim3(1), bv, 0, st, pa.p, pb.p, tol, S.p);
    BicgScalars h{};
    MG_HIP(hipMemcpyAsync(&h, S.p, sizeof(h), hipMemcpyDeviceToHost, st));
    MG_HIP(hipStreamSynchronize(st));
    if (!h.done) {
        // the preconditioner is built only once the initial residual has not passed (:78-83)
        const Ilu0 *M = pre;
        if (!M && use_pre) {
            owned.reset(new Ilu0());
            ilu0_build(*owned, A, btol, st);
            M = owned.get();
        }
        auto apply = [&](const double *src, double *dst) {
            if (M) ilu0_apply_device(*M, src, dst, st, done);
            else hipLaunchKernelGGL(k_bg_copy, gv, bv, 0, st, src, dst, n, S.p);
        };
        const int maxit = max_it > 0 ? max_it : n;
        for (int it = 1; it <= maxit; ++it) {
            hipLaunchKernelGGL(k_bg_dot, gr, bv, 0, st, rh.p, r.p, n, pa.p, S.p);
            hipLaunchKernelGGL(k_bg_rho, dim3(1), bv, 0, st, pa.p, it, btol, S.p);
            hipLaunchKernelGGL(k_bg_direction, gv, bv, 0, st, r.p, p.p, v.p, n, it, S.p);
            apply(p.p, ph.p);
            hipLaunchKernelGGL(k_gspmv, gv, bv, 0, st, A.rowptr.p, A.colidx.p, A.vals.p, n, ph.p, v.p, done);
            hipLaunchKernelGGL(k_bg_dot, gr, bv, 0, st, rh.p, v.p, n, pa.p, S.p);
            hipLaunchKernelGGL(k_bg_alpha, dim3(1), bv, 0, st, pa.p, it, btol, S.p);
            hipLaunchKernelGGL(k_bg_s, gr, bv, 0, st, r.p, v.p, s.p, n, pa.p, S.p);
            hipLaunchKernelGGL(k_bg_snorm, dim3(1), bv, 0, st, pa.p, S.p);
            hipLaunchKernelGGL(k_bg_xearly, gv, bv, 0, st, d_x, ph.p, n, S.p);
            hipLaunchKernelGGL(k_bg_early_done, dim3(1), dim3(64), 0, st, it, S.p);
            apply(s.p, sh.p);
            hipLaunchKernelGGL(k_gspmv, gv, bv, 0, st, A.rowptr.p, A.colidx.p, A.vals.p, n, sh.p, t.p, done);
            hipLaunchKernelGGL(k_bg_dot, gr, bv, 0, st, t.p, t.p, n, pa.p, S.p);
            hipLaunchKernelGGL(k_bg_dot, gr, bv, 0, st, t.p, s.p, n, pb.p, S.p);
            hipLaunchKernelGGL(k_bg_omega, dim3(1), bv, 0, st, pa.p, pb.p, it, btol, S.p);
            hipLaunchKernelGGL(k_bg_update, gr, bv, 0, st, d_x, r.p, ph.p, sh.p, s.p, t.p, n, pa.p, S.p);
            hipLaunchKernelGGL(k_bg_rnorm, dim3(1), bv, 0, st, pa.p, it, S.p);
            if (it % 8 == 0 || it == maxit) {
                MG_HIP(hipMemcpyAsync(&h, S.p, sizeof(h), hipMemcpyDeviceToHost, st));
                MG_HIP(hipStreamSynchronize(st));
                if (h.done) break;
            }
        }
        MG_HIP(hipGetLastError());
        MG_HIP(hipMemcpyAsync(&h, S.p, sizeof(h), hipMemcpyDeviceToHost, st));
        MG_HIP(hipStreamSynchronize(st));
        if (h.fail) {
            static const char *what[] = {"", "rho", "omega", "(r0, A M^-1 p)", "(A M^-1 s, A M^-1 s)", "omega"};
            throw Error(MG_E_NOT_PD, std::string("BiCGStab breakdown: ") + what[h.fail] +
                                         " is zero or non-finite at iteration " + std::to_string(h.fail_it) + ".");
        }
        if (!h.done) h.iters = maxit;
    }
    res.iterations = h.iters;
    res.residual_norm = h.rnorm;
    res.relative_residual = h.rnorm / h.bnorm;
    res.converged = res.relative_residual <= tol;
}

}  // namespace mg
