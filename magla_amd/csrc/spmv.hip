// spmv.hip -- symmetric SpMV and the Jacobi-PCG loop (CGGP) on MI355X.
//
// The upper-triangle SymmetricCSRMatrix is expanded once into a full CSR
// (both triangles, columns ascending) so that y = A x is a row-parallel
// gather with no atomics: a group of `lanes` lanes owns one row, partial sums
// are combined by a fixed xor-shuffle tree (deterministic).  Algorithmic
// traffic per product: 12 B per stored entry + 4 B per row pointer + 16 B per
// row and right-hand side (x read once, y written once).
//
// PCG keeps every vector and scalar in HBM.  An iteration is two kernels: the new
// direction p = z + beta p formed inside the SpMV's gather (into the other of two
// direction buffers) with Ap and p'Ap, then the x / r update with z = M r, ||r||^2 and
// r'z.  Each reduction is finished by the workgroup that arrives last, in a fixed order
// (deterministic), which also computes alpha, or the convergence test and beta.  A device
// flag turns the remaining launches into no-ops once ||r|| <= tol*||b|| (same test, same
// place as ConjugateGradientGaussPreSolver.cs:151-153).
#include <algorithm>
#include <cmath>
#include <map>
#include <memory>

#include "device.h"
#include "reduce.h"

namespace mg {

namespace {
constexpr int kBS = 256;

inline unsigned nblk(int64_t count, int bs) { return (unsigned)((count + bs - 1) / bs); }
}  // namespace

__global__ void k_full_gather(const int *__restrict__ src, int64_t nnz, const double *__restrict__ upper,
                              double *__restrict__ full) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nnz) full[e] = upper[src[e]];
}

// Jacobi inverse diagonal (JacobiPreconditioner.cs:24-31); status[0] = 1 on
// a zero / non-finite diagonal.
__global__ void k_inv_diag(const int *__restrict__ rp, const int *__restrict__ ci, const double *__restrict__ v, int n,
                           double *__restrict__ invd, int *__restrict__ status) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k)
        if (ci[k] == i) d = v[k];
    if (!isfinite(d) || fabs(d) <= 1e-300) {
        atomicOr(status, 1);
        invd[i] = 0.0;
    } else {
        invd[i] = 1.0 / d;
    }
}

template <int LANES>
__global__ __launch_bounds__(256) void k_spmv(const int *__restrict__ rp, const int *__restrict__ ci,
                                              const double *__restrict__ v, int n, const double *__restrict__ x,
                                              double *__restrict__ y) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = gid / LANES, sub = gid % LANES;
    if (row >= n) return;
    double acc = 0.0;
    for (int k = rp[row] + sub; k < rp[row + 1]; k += LANES) acc += v[k] * x[ci[k]];
    for (int o = LANES / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, LANES);
    if (sub == 0) y[row] = acc;
}

void Matrix::build(int n_, const int *rowptr, const int *colidx) {
    n = n_;
    nnz_upper = n ? rowptr[n] : 0;
    h_rowptr.assign(rowptr, rowptr + n + 1);
    h_colidx.assign(colidx, colidx + nnz_upper);
    std::vector<int> frp(n + 1, 0);
    for (int r = 0; r < n; ++r)
        for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) {
            ++frp[r + 1];
            if (colidx[k] != r) ++frp[colidx[k] + 1];
        }
    for (int i = 0; i < n; ++i) frp[i + 1] += frp[i];
    nnz_full = frp[n];
    std::vector<int> fci(nnz_full), fsrc(nnz_full), at(frp.begin(), frp.end() - 1);
    for (int r = 0; r < n; ++r)
        for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) {
            const int c = colidx[k];
            fci[at[r]] = c;
            fsrc[at[r]++] = k;
            if (c != r) {
                fci[at[c]] = r;
                fsrc[at[c]++] = k;
            }
        }
    // row [lower entries from earlier rows (ascending), own upper entries] is sorted;
    // rows are filled in ascending r, so the lower part arrives ascending.
    std::vector<int> dp(n, -1);
    for (int i = 0; i < n; ++i)
        for (int k = frp[i]; k < frp[i + 1]; ++k)
            if (fci[k] == i) dp[i] = k;
    this->rowptr.upload(frp);
    if (fci.empty()) fci.push_back(0);  // at least one entry: kernels may read a clamped index
    this->colidx.upload(fci);
    src.upload(fsrc);
    dpos.upload(dp);
    vals.alloc(std::max<int64_t>(nnz_full, 1));
    invdiag.alloc(std::max(n, 1));
    // sliced ELL: one 64-row slice per wave, lane = row, the slice padded to its longest row
    nslices = (n + 63) / 64;
    std::vector<int> sp(nslices + 1, 0);
    for (int sl = 0; sl < nslices; ++sl) {
        int w = 0;
        for (int r = 64 * sl; r < std::min(n, 64 * sl + 64); ++r) w = std::max(w, frp[r + 1] - frp[r]);
        sp[sl + 1] = sp[sl] + 64 * w;
    }
    sell = n > 0 && (double)sp[nslices] <= 1.5 * (double)nnz_full;
    if (sell) {
        std::vector<int> sci(sp[nslices]), ssrc(sp[nslices]);
        for (int sl = 0; sl < nslices; ++sl) {
            const int w = (sp[sl + 1] - sp[sl]) / 64;
            for (int l = 0; l < 64; ++l) {
                const int r = std::min(64 * sl + l, n - 1);
                const int len = 64 * sl + l < n ? frp[r + 1] - frp[r] : 0;
                for (int k = 0; k < w; ++k) {
                    const int e = sp[sl] + 64 * k + l;
                    sci[e] = k < len ? fci[frp[r] + k] : r;
                    ssrc[e] = k < len ? fsrc[frp[r] + k] : -1;
                }
            }
        }
        sell_ptr.upload(sp);
        sell_ci.upload(sci);
        sell_src.upload(ssrc);
        sell_vals.alloc(std::max(sp[nslices], 1));
    }
    const double avg = n ? (double)nnz_full / n : 1.0;
    // About ten entries per lane: a wave then works on several rows at once, which
    // keeps enough independent loads in flight (one row per wave is latency-bound).
    lanes = avg > 512 ? 64 : avg > 256 ? 32 : avg > 128 ? 16 : avg > 24 ? 8 : 4;
}

__global__ void k_sell_gather(const int *__restrict__ src, int64_t count, const double *__restrict__ upper,
                              double *__restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < count) out[e] = src[e] >= 0 ? upper[src[e]] : 0.0;
}

void Matrix::set_values_dev(const double *d_values, hipStream_t st) {
    if (nnz_full > 0)
        hipLaunchKernelGGL(k_full_gather, dim3(nblk(nnz_full, kBS)), dim3(kBS), 0, st, src.p, nnz_full, d_values,
                           vals.p);
    if (sell && sell_src.n > 0)
        hipLaunchKernelGGL(k_sell_gather, dim3(nblk((int64_t)sell_src.n, kBS)), dim3(kBS), 0, st, sell_src.p,
                           (int64_t)sell_src.n, d_values, sell_vals.p);
    MG_HIP(hipGetLastError());
}

void Matrix::set_values_host(const double *values) {
    DevArray<double> tmp(std::max(nnz_upper, 1));
    if (nnz_upper) MG_HIP(hipMemcpy(tmp.p, values, sizeof(double) * nnz_upper, hipMemcpyHostToDevice));
    set_values_dev(tmp.p, 0);
    MG_HIP(hipDeviceSynchronize());
    DevArray<int> status(1);
    MG_HIP(hipMemset(status.p, 0, sizeof(int)));
    if (n > 0)
        hipLaunchKernelGGL(k_inv_diag, dim3(nblk(n, kBS)), dim3(kBS), 0, 0, rowptr.p, colidx.p, vals.p, n, invdiag.p,
                           status.p);
    int h = 0;
    MG_HIP(hipMemcpy(&h, status.p, sizeof(int), hipMemcpyDeviceToHost));
    diag_status = h ? MG_E_NOT_PD : 0;
}

static void launch_spmv(const Matrix &A, const double *x, double *y, hipStream_t st) {
    const int64_t threads = (int64_t)A.n * A.lanes;
    const dim3 g(nblk(threads, kBS)), b(kBS);
    switch (A.lanes) {
        case 4: hipLaunchKernelGGL(k_spmv<4>, g, b, 0, st, A.rowptr.p, A.colidx.p, A.vals.p, A.n, x, y); break;
        case 8: hipLaunchKernelGGL(k_spmv<8>, g, b, 0, st, A.rowptr.p, A.colidx.p, A.vals.p, A.n, x, y); break;
        case 16: hipLaunchKernelGGL(k_spmv<16>, g, b, 0, st, A.rowptr.p, A.colidx.p, A.vals.p, A.n, x, y); break;
        case 32: hipLaunchKernelGGL(k_spmv<32>, g, b, 0, st, A.rowptr.p, A.colidx.p, A.vals.p, A.n, x, y); break;
        default: hipLaunchKernelGGL(k_spmv<64>, g, b, 0, st, A.rowptr.p, A.colidx.p, A.vals.p, A.n, x, y); break;
    }
}

void spmv_device(const Matrix &A, int nrhs, const double *d_x, int ldx, double *d_y, int ldy, hipStream_t st) {
    if (A.n == 0) return;
    for (int c = 0; c < nrhs; ++c) launch_spmv(A, d_x + (int64_t)c * ldx, d_y + (int64_t)c * ldy, st);
    MG_HIP(hipGetLastError());
}

// ---- PCG ---------------------------------------------------------------------


__device__ inline double block_sum(double v, double *sm) {
    sm[threadIdx.x] = v;
    __syncthreads();
    for (int h = kBS / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sm[threadIdx.x] += sm[threadIdx.x + h];
        __syncthreads();
    }
    return sm[0];
}

// partial[b] = sum_i a_i * c_i over a grid-stride range (fixed grid)
__global__ __launch_bounds__(256) void k_dot_part(const double *__restrict__ a, const double *__restrict__ c, int n,
                                                  double *__restrict__ part, const PcgScalars *__restrict__ S) {
    __shared__ double sm[kBS];
    if (S && S->done) return;
    double acc = 0.0;
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) acc += a[i] * c[i];
    const double t = block_sum(acc, sm);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__device__ inline double finish_sum_n(const double *part, int nparts, double *sm) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < nparts; b += kBS) acc += part[b];
    return block_sum(acc, sm);
}

__device__ inline double finish_sum(const double *part, double *sm) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < kRedBlocks; b += kBS) acc += part[b];
    return block_sum(acc, sm);
}

// r = b - A x0 already computed in r; s->rnorm, threshold and early exit
__global__ void k_pcg_init(const double *__restrict__ part_rr, const double *__restrict__ part_bb, double tol,
                           PcgScalars *__restrict__ S) {
    __shared__ double sm[kBS];
    const double rr = finish_sum(part_rr, sm);
    __syncthreads();
    const double bb = finish_sum(part_bb, sm);
    if (threadIdx.x == 0) {
        const double bn = sqrt(bb);
        const double scale = bn > 0.0 ? bn : 1.0;
        S->bnorm = scale;
        S->thr = tol * scale;
        S->rnorm = sqrt(rr);
        S->done = S->rnorm <= S->thr ? 1 : 0;
        S->iters = 0;
        S->fail = 0;
    }
}

__global__ void k_residual0(const double *__restrict__ b, const double *__restrict__ ax, int n, double *__restrict__ r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = b[i] - ax[i];
}

__global__ void k_precond(const double *__restrict__ invd, const double *__restrict__ r, int n, double *__restrict__ z,
                          int use, const PcgScalars *__restrict__ S) {
    if (S->done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = use ? invd[i] * r[i] : r[i];
}

template <int LANES>
__global__ __launch_bounds__(256) void k_spmv_guard(const int *__restrict__ rp, const int *__restrict__ ci,
                                                    const double *__restrict__ v, int n, const double *__restrict__ x,
                                                    double *__restrict__ y, const PcgScalars *__restrict__ S) {
    if (S->done) return;
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = gid / LANES, sub = gid % LANES;
    if (row >= n) return;
    double acc = 0.0;
    for (int k = rp[row] + sub; k < rp[row + 1]; k += LANES) acc += v[k] * x[ci[k]];
    for (int o = LANES / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, LANES);
    if (sub == 0) y[row] = acc;
}

// alpha = rz / p'Ap  (ConjugateGradientGaussPreSolver.cs:137-143)
__global__ void k_pcg_alpha(const double *__restrict__ part, PcgScalars *__restrict__ S) {
    __shared__ double sm[kBS];
    if (S->done) return;
    const double pap = finish_sum(part, sm);
    if (threadIdx.x == 0) {
        S->pap = pap;
        if (pap <= 0.0) {
            S->fail = 1;
            S->done = 1;
        }
        S->alpha = S->rz / pap;
    }
}

// x += alpha p ; r -= alpha Ap ; partial ||r||^2  (:145-151)
__global__ __launch_bounds__(256) void k_pcg_update(double *__restrict__ x, double *__restrict__ r,
                                                    const double *__restrict__ p, const double *__restrict__ ap, int n,
                                                    double *__restrict__ part, const PcgScalars *__restrict__ S) {
    __shared__ double sm[kBS];
    if (S->done) return;
    const double alpha = S->alpha;
    double acc = 0.0;
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * ap[i];
        r[i] = ri;
        acc += ri * ri;
    }
    const double t = block_sum(acc, sm);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// convergence test (:151-153)
__global__ void k_pcg_norm(const double *__restrict__ part, int it, PcgScalars *__restrict__ S) {
    __shared__ double sm[kBS];
    if (S->done) return;
    const double rr = finish_sum(part, sm);
    if (threadIdx.x == 0) {
        S->rnorm = sqrt(rr);
        S->iters = it;
        if (S->rnorm <= S->thr) S->done = 1;
    }
}

// beta = rz_new / rz (:157), rz = rz_new (:162)
__global__ void k_pcg_beta(const double *__restrict__ part, PcgScalars *__restrict__ S) {
    __shared__ double sm[kBS];
    if (S->done) return;
    const double rz = finish_sum(part, sm);
    if (threadIdx.x == 0) {
        S->beta = rz / S->rz;
        S->rz = rz;
    }
}

// p = z + beta p (:159-160)
__global__ void k_pcg_pupdate(const double *__restrict__ z, double *__restrict__ p, int n,
                              const PcgScalars *__restrict__ S) {
    if (S->done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = z[i] + S->beta * p[i];
}

// S->done as the workgroup's first thread reads it: workgroup 0 of the same kernel may set it
// meanwhile, and every thread of a workgroup must take the same branch.
__device__ __forceinline__ bool pcg_done(const PcgScalars *__restrict__ S) {
    __shared__ int done;
    if (threadIdx.x == 0) done = S->done;
    __syncthreads();
    return done != 0;
}

// the direction entry p_new = z + beta p_old, formed identically by its row's owner and by
// every row that gathers it
__device__ __forceinline__ double pcg_dir(double z, double p, double beta) { return __fma_rn(beta, p, z); }

// Flexible PCG (the learned preconditioner), iteration `it`, first kernel on the sliced ELL copy
// (wave = 64-row slice, lane = row): p_new = z + beta p_old in the gather (beta from the
// previous iteration's partials; beta = 0 and p_old = 0 give the first direction p = z),
// Ap = A p_new and partial p'Ap.  The first kPre entries of each row are loaded before the
// partial sums are formed.  The convergence test is the preconditioner kernels'.
__global__ __launch_bounds__(256) void k_pcg_dir_sell(const int *__restrict__ sp, const int *__restrict__ sci,
                                                      const double *__restrict__ sv, int n, int nslices,
                                                      const double *__restrict__ z, const double *__restrict__ p_old,
                                                      double *__restrict__ p_new, double *__restrict__ ap,
                                                      const double *__restrict__ part_rr,
                                                      const double *__restrict__ part_rz, int np_step, int it,
                                                      double *__restrict__ part_pap, PcgScalars *__restrict__ S,
                                                      FcgState *__restrict__ fsp) {
    __shared__ double sm4[4];
    if (pcg_done(S)) return;  // an iteration queued past convergence
    constexpr int kPre = 8;
    const int lane = threadIdx.x & 63;
    int slice = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    const int s0 = min(slice, nslices - 1);
    const int base0 = sp[s0], w0 = slice < nslices ? (sp[s0 + 1] - base0) >> 6 : 0;
    const int r0 = min(64 * s0 + lane, n - 1);
    // Every prefetch load is unconditional (entries past the slice's width repeat its last
    // one and are zeroed where they are used): a load under a branch lands in a temporary
    // whose copy waits for it, and the waits would serialize the prefetch.
    int cpre[kPre];
    double vpre[kPre], zpre[kPre], ppre[kPre];
#pragma unroll
    for (int k = 0; k < kPre; ++k) {
        const int e = base0 + 64 * min(k, max(w0 - 1, 0)) + lane;
        cpre[k] = sci[e];
        vpre[k] = sv[e];
    }
#pragma unroll
    for (int k = 0; k < kPre; ++k) {
        zpre[k] = z[cpre[k]];
        ppre[k] = p_old[cpre[k]];
    }
    const double zr = z[r0], pr = p_old[r0];
    double beta = 0.0;
    {
        // the flexible direction update of iteration it - 1 (part_rr / part_rz: its r'z and
        // z'Ap partials): beta = -(z'Ap) / (p'Ap), or 0 on a restart or a non-positive r'z
        // (which also falls back to Jacobi); workgroup 0 completes the state slot
        if (it > 1) {
            const double rz = sum_partials(part_rr, np_step, sm4), zap = sum_partials(part_rz, np_step, sm4);
            int fb = fsp->fallback1, rs = fsp->restart1;
            if (rs || !(rz > 0.0)) {
                beta = 0.0;
                const int fb0 = fb;
                rs = 0;
                if (!(rz > 0.0) && !fb0) {
                    fb = 1;
                    rs = 1;
                }
            } else {
                beta = -zap / S->pap;
            }
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                fsp->rz2 = rz;
                fsp->beta = beta;
                fsp->fallback2 = fb;
                fsp->restart2 = rs;
                S->fallback = fb;
            }
        }
    }
    double dot = 0.0;
    for (bool first = true; slice < nslices; slice += gridDim.x * 4, first = false) {
        const int row = 64 * slice + lane;
        const int base = first ? base0 : sp[slice], w = first ? w0 : (sp[slice + 1] - base) >> 6;
        double acc = 0.0;
        int k = 0;
        if (first) {
#pragma unroll
            for (int q = 0; q < kPre; ++q) acc += (q < w0 ? vpre[q] : 0.0) * pcg_dir(zpre[q], ppre[q], beta);
            k = kPre;
        }
        for (; k < w; ++k) {
            const int e = base + 64 * k + lane;
            const int c = sci[e];
            acc += sv[e] * pcg_dir(z[c], p_old[c], beta);
        }
        if (row < n) {
            const double pn = first ? pcg_dir(zr, pr, beta) : pcg_dir(z[row], p_old[row], beta);
            p_new[row] = pn;
            ap[row] = acc;
            dot += pn * acc;
        }
    }
    const double t = block_sum_waves(dot, sm4);
    if (threadIdx.x == 0) part_pap[blockIdx.x] = t;
}

// Flexible PCG, iteration `it`, second kernel: alpha = rz / p'Ap (:137-143; p'Ap <= 0 fails;
// rz from *rzp), x += alpha p, r -= alpha Ap (:145-150), partial ||r||^2; p'Ap and alpha are
// kept in S.  Each thread's first kPre elements are loaded before the partial sum of p'Ap is
// formed.
__global__ __launch_bounds__(256) void k_pcg_step(double *__restrict__ x, double *__restrict__ r,
                                                  const double *__restrict__ p, const double *__restrict__ ap, int n,
                                                  const double *__restrict__ part_pap, int np_dir,
                                                  double *__restrict__ part_rr, PcgScalars *__restrict__ S,
                                                  const double *__restrict__ rzp) {
    __shared__ double sm4[4];
    constexpr int kPre = 4;
    const int stride = gridDim.x * kBS, i0 = blockIdx.x * kBS + threadIdx.x;
    double px[kPre], pp[kPre], pr[kPre], pa[kPre];
#pragma unroll
    for (int e = 0; e < kPre; ++e) {
        const int i = min(i0 + e * stride, n - 1);
        px[e] = x[i];
        pp[e] = p[i];
        pr[e] = r[i];
        pa[e] = ap[i];
    }
    const double pap = sum_partials(part_pap, np_dir, sm4);
    if (S->done) return;  // uniform
    if (pap <= 0.0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            S->pap = pap;
            S->fail = 1;
            S->done = 1;
        }
        return;
    }
    const double alpha = *rzp / pap;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        S->pap = pap;
        S->alpha = alpha;
    }
    double rr = 0.0;
#pragma unroll
    for (int e = 0; e < kPre; ++e) {
        const int i = i0 + e * stride;
        if (i < n) {
            x[i] = px[e] + alpha * pp[e];
            const double ri = pr[e] - alpha * pa[e];
            r[i] = ri;
            rr += ri * ri;
        }
    }
    for (int i = i0 + kPre * stride; i < n; i += stride) {
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * ap[i];
        r[i] = ri;
        rr += ri * ri;
    }
    const double t0 = block_sum_waves(rr, sm4);
    if (threadIdx.x == 0) part_rr[blockIdx.x] = t0;
}

// ---- single-reduction Jacobi-PCG (one kernel per iteration) --------------------------
//
// ConjugateGradientGaussPreSolver's PCG (CGGP/ConjugateGradientGaussPreSolver.cs:78-166) in the
// Chronopoulos-Gear form: the same iterates in exact arithmetic, with one reduction per
// iteration.  With u = M r and w = A u, gamma = r'u and delta = w'u give both scalars:
// beta_i = gamma_i / gamma_{i-1}, and p_i'A p_i = delta_i - beta_i gamma_i / alpha_{i-1}
// (the reference's p'Ap <= 0 failure is tested on it), alpha_i = gamma_i / p_i'A p_i.  The
// directions carry s = A p by the same recurrence as p (p_i = u_i + beta_i p_{i-1},
// s_i = w_i + beta_i s_{i-1}), so x_{i+1} = x_i + alpha_i p_i and r_{i+1} = r_i - alpha_i s_i
// need no product with A of their own.  Kernel i updates its rows and forms w_{i+1} = A u_{i+1}:
// each neighbour's u_{i+1} is recomputed from iteration i's r, w and s (which alternate between
// two buffers, so no row is overwritten while another workgroup reads it), by the same
// operations as the neighbour's own row, so both agree bit for bit.  Its partials of gamma_{i+1},
// delta_{i+1} and ||r_{i+1}||^2 go to the next kernel, whose every workgroup sums them in a fixed
// order; that kernel also makes iteration i + 1's convergence test (:151-153).
struct Cg1Vec {
    const double *r, *w, *s;  // iteration i: r_i, w_i, s_{i-1}
    double *r1, *w1, *s1;     // written: r_{i+1}, w_{i+1}, s_i
};
__device__ __forceinline__ double cg1_u(double r, double invd, int use) { return use ? invd * r : r; }
// u_{i+1} of row c from iteration i's values
__device__ __forceinline__ double cg1_u_next(double r, double w, double s, double invd, double alpha, double beta,
                                             int use) {
    return cg1_u(__fma_rn(-alpha, __fma_rn(beta, s, w), r), invd, use);
}

// Block sums of three values at once (one pass of wave reductions, two barriers), in a fixed
// order; sm12 holds 12 doubles.
__device__ __forceinline__ void block_sum3(double &a, double &b, double &c, double *sm12) {
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
        c += __shfl_xor(c, o, 64);
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sm12[wv] = a;
        sm12[4 + wv] = b;
        sm12[8 + wv] = c;
    }
    __syncthreads();
    a = (sm12[0] + sm12[1]) + (sm12[2] + sm12[3]);
    b = (sm12[4] + sm12[5]) + (sm12[6] + sm12[7]);
    c = (sm12[8] + sm12[9]) + (sm12[10] + sm12[11]);
    __syncthreads();
}
// sums of the three partial arrays part[0..np), [np..2np), [2np..3np) in one pass
__device__ __forceinline__ void sum_partials3(const double *__restrict__ part, int np, double *sm12, double &a,
                                              double &b, double &c) {
    a = b = c = 0.0;
    for (int i = threadIdx.x; i < np; i += kBS) {
        a += part[i];
        b += part[np + i];
        c += part[2 * np + i];
    }
    block_sum3(a, b, c, sm12);
}

// Scalars of kernel `it` (iteration i = it - 1; it = 0 is the start pass, w_0 = A M r_0, with
// alpha = beta = 0) from the previous kernel's partials part[0..np) (gamma), [np..2np) (delta),
// [2np..3np) (||r||^2).  False when the kernel does nothing: done before, iteration it - 1
// converged, or p'Ap <= 0.  Workgroup 0 records the state.
__device__ bool cg1_scalars(const double *__restrict__ part, int np, int it, PcgScalars *__restrict__ S,
                            double *sm12, double &alpha, double &beta) {
    alpha = beta = 0.0;
    if (it == 0) return true;
    double gamma, delta, rho;
    sum_partials3(part, np, sm12, gamma, delta, rho);
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    double pap = delta;  // p_0 = u_0
    if (it > 1) {
        const double rnorm = sqrt(rho);
        if (rnorm <= S->thr) {
            if (lead) {
                S->rnorm = rnorm;
                S->iters = it - 1;
                S->done = 1;
            }
            return false;
        }
        if (lead) {
            S->rnorm = rnorm;
            S->iters = it - 1;
        }
        beta = gamma / S->rz2[(it - 1) & 1];
        pap = delta - beta * gamma / S->al2[(it - 1) & 1];
    }
    if (pap <= 0.0) {
        if (lead) {
            S->pap = pap;
            S->fail = 1;
            S->done = 1;
        }
        return false;
    }
    alpha = gamma / pap;
    if (lead) {
        S->rz2[it & 1] = gamma;
        S->al2[it & 1] = alpha;
    }
    return true;
}

// A row's own update; returns its u_{i+1} and adds its terms to the partials.
// the row's own values of iteration i
struct Cg1Own {
    double r, w, s, d, p, x;
};
__device__ __forceinline__ Cg1Own cg1_own(int row, const Cg1Vec &v, const double *__restrict__ invd, int use,
                                          const double *__restrict__ x, const double *__restrict__ p) {
    return Cg1Own{v.r[row], v.w[row], v.s[row], use ? invd[row] : 1.0, p[row], x[row]};
}
__device__ __forceinline__ void cg1_row(int row, double aw, const Cg1Own &o, const Cg1Vec &v, int use,
                                        double *__restrict__ x, double *__restrict__ p, double alpha, double beta,
                                        double &g, double &dl, double &rr) {
    const double r = o.r, d = o.d;
    const double sn = __fma_rn(beta, o.s, o.w);
    const double pn = __fma_rn(beta, o.p, cg1_u(r, d, use));
    const double r1 = __fma_rn(-alpha, sn, r);
    const double u1 = cg1_u(r1, d, use);
    x[row] = __fma_rn(alpha, pn, o.x);
    p[row] = pn;
    v.s1[row] = sn;
    v.r1[row] = r1;
    v.w1[row] = aw;
    g += r1 * u1;
    dl += aw * u1;
    rr += r1 * r1;
}

__device__ __forceinline__ void cg1_partials(double g, double dl, double rr, double *__restrict__ part_out,
                                             double *sm12) {
    block_sum3(g, dl, rr, sm12);
    if (threadIdx.x == 0) {
        part_out[blockIdx.x] = g;
        part_out[gridDim.x + blockIdx.x] = dl;
        part_out[2 * gridDim.x + blockIdx.x] = rr;
    }
}

// sliced ELL: wave = 64-row slice, lane = row
__global__ __launch_bounds__(256) void k_pcg_cg1_sell(const int *__restrict__ sp, const int *__restrict__ sci,
                                                      const double *__restrict__ sv, int n, int nslices,
                                                      const double *__restrict__ invd, int use, double *__restrict__ x,
                                                      double *__restrict__ p, Cg1Vec v,
                                                      const double *__restrict__ part_in, int np_in,
                                                      double *__restrict__ part_out, int it,
                                                      PcgScalars *__restrict__ S) {
    __shared__ double sm12[12];
    if (pcg_done(S)) return;  // uniform
    // The first slice's first kPre entries and the values they gather are loaded before the
    // scalars are summed (none of them depends on alpha or beta), so their latency overlaps it.
    constexpr int kPre = 8;
    const int lane = threadIdx.x & 63;
    int slice = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    const int s0 = min(slice, nslices - 1);
    const int base0 = sp[s0], w0 = slice < nslices ? (sp[s0 + 1] - base0) >> 6 : 0;
    // All column indices first, then every value and gather, each load unconditional (entries
    // past the slice's width repeat its last one and are zeroed where they are used; without
    // the Jacobi scaling the D^-1 slot reads r): a load under a branch lands in a temporary
    // whose copy waits for it, and in-order waits would serialize the prefetch entry by entry.
    int cpre[kPre];
    double vpre[kPre], rpre[kPre], wpre[kPre], spre[kPre], dpre[kPre];
    const double *dsrc = use ? invd : v.r;
#pragma unroll
    for (int k = 0; k < kPre; ++k) cpre[k] = sci[base0 + 64 * min(k, max(w0 - 1, 0)) + lane];
#pragma unroll
    for (int k = 0; k < kPre; ++k) {
        const int c = cpre[k];
        vpre[k] = sv[base0 + 64 * min(k, max(w0 - 1, 0)) + lane];
        rpre[k] = v.r[c];
        wpre[k] = v.w[c];
        spre[k] = v.s[c];
        dpre[k] = dsrc[c];
    }
    double alpha, beta;
    if (!cg1_scalars(part_in, np_in, it, S, sm12, alpha, beta)) return;  // uniform
    double g = 0.0, dl = 0.0, rr = 0.0;
    for (bool first = true; slice < nslices; slice += gridDim.x * 4, first = false) {
        const int row = 64 * slice + lane;
        const int base = first ? base0 : sp[slice], w = first ? w0 : (sp[slice + 1] - base) >> 6;
        double acc = 0.0;
        int k = 0;
        if (first) {
#pragma unroll
            for (int q = 0; q < kPre; ++q)
                acc += (q < w0 ? vpre[q] : 0.0) *
                       cg1_u_next(rpre[q], wpre[q], spre[q], use ? dpre[q] : 1.0, alpha, beta, use);
            k = kPre;
        }
        for (; k < w; ++k) {
            const int e = base + 64 * k + lane;
            const int c = sci[e];
            acc += sv[e] * cg1_u_next(v.r[c], v.w[c], v.s[c], use ? invd[c] : 1.0, alpha, beta, use);
        }
        if (row < n) cg1_row(row, acc, cg1_own(row, v, invd, use, x, p), v, use, x, p, alpha, beta, g, dl, rr);
    }
    cg1_partials(g, dl, rr, part_out, sm12);
}

// CSR: LANES lanes per row
template <int LANES>
__global__ __launch_bounds__(256) void k_pcg_cg1_csr(const int *__restrict__ rp, const int *__restrict__ ci,
                                                     const double *__restrict__ val, int n,
                                                     const double *__restrict__ invd, int use, double *__restrict__ x,
                                                     double *__restrict__ p, Cg1Vec v,
                                                     const double *__restrict__ part_in, int np_in,
                                                     double *__restrict__ part_out, int it,
                                                     PcgScalars *__restrict__ S) {
    __shared__ double sm12[12];
    if (pcg_done(S)) return;  // uniform
    double alpha, beta;
    if (!cg1_scalars(part_in, np_in, it, S, sm12, alpha, beta)) return;  // uniform
    constexpr int kRows = kBS / LANES;
    const int sub = threadIdx.x % LANES;
    double g = 0.0, dl = 0.0, rr = 0.0;
    for (int row = blockIdx.x * kRows + (int)threadIdx.x / LANES; row < n; row += gridDim.x * kRows) {
        double acc = 0.0;
        for (int k = rp[row] + sub; k < rp[row + 1]; k += LANES) {
            const int c = ci[k];
            acc += val[k] * cg1_u_next(v.r[c], v.w[c], v.s[c], use ? invd[c] : 1.0, alpha, beta, use);
        }
        for (int o = LANES / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, LANES);
        if (sub == 0) cg1_row(row, acc, cg1_own(row, v, invd, use, x, p), v, use, x, p, alpha, beta, g, dl, rr);
    }
    cg1_partials(g, dl, rr, part_out, sm12);
}

// after the last iteration: its convergence test
__global__ void k_pcg_cg1_last(const double *__restrict__ part, int np, int it, PcgScalars *__restrict__ S) {
    __shared__ double sm12[12];
    if (pcg_done(S)) return;
    double gamma, delta, rho;
    sum_partials3(part, np, sm12, gamma, delta, rho);
    const double rnorm = sqrt(rho);
    if (threadIdx.x == 0) {
        S->rnorm = rnorm;
        S->iters = it;
        if (rnorm <= S->thr) S->done = 1;
    }
}

// Host polling of the iteration state without draining the stream: batch k of iterations ends
// with a copy of the scalars into pinned slot k & 1 and an event; the host then waits for batch
// k - 1's event only, so batch k keeps the GPU busy meanwhile.  One per host thread and device
// (pinned memory and events are made once).
struct ScalarPoll {
    PcgScalars *h = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int pending = -1;
    ScalarPoll() {
        MG_HIP(hipHostMalloc(reinterpret_cast<void **>(&h), 2 * sizeof(PcgScalars), hipHostMallocDefault));
        for (hipEvent_t &e : ev) MG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    ~ScalarPoll() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        if (h) (void)hipHostFree(h);
    }
    static ScalarPoll &get() {
        static thread_local std::map<int, std::unique_ptr<ScalarPoll>> polls;
        int dev = 0;
        MG_HIP(hipGetDevice(&dev));
        auto &q = polls[dev];
        if (!q) q.reset(new ScalarPoll());
        q->pending = -1;
        return *q;
    }
    // enqueues batch `batch`'s copy; true when the previous batch had ended with done set
    bool post(const PcgScalars *S, int batch, hipStream_t st) {
        const int slot = batch & 1;
        MG_HIP(hipMemcpyAsync(h + slot, S, sizeof(PcgScalars), hipMemcpyDeviceToHost, st));
        MG_HIP(hipEventRecord(ev[slot], st));
        const int prev = pending;
        pending = slot;
        if (prev < 0) return false;
        MG_HIP(hipEventSynchronize(ev[prev]));
        return h[prev].done != 0;
    }
};

// grids (fixed per matrix, so the reductions are deterministic): the direction kernels cover
// the rows in one pass up to 4096 workgroups; the flexible PCG's step kernel's workgroups leave
// their partials for the next direction kernel's workgroups to sum
static int pcg_dir_blocks(const Matrix &A) {
    return A.sell ? std::min(1024, (A.nslices + 3) / 4) : (int)std::min<int64_t>(1024, nblk((int64_t)A.n * A.lanes, kBS));
}
static int pcg_step_blocks(int n) { return (int)std::min<int64_t>(512, nblk(n, kBS)); }

static void launch_spmv_guard(const Matrix &A, const double *x, double *y, const PcgScalars *S, hipStream_t st) {
    const int64_t threads = (int64_t)A.n * A.lanes;
    const dim3 g(nblk(threads, kBS)), b(kBS);
    switch (A.lanes) {
        case 4: hipLaunchKernelGGL(k_spmv_guard<4>, g, b, 0, st, A.rowptr.p, A.colidx.p, A.vals.p, A.n, x, y, S); break;
        case 8: hipLaunchKernelGGL(k_spmv_guard<8>, g, b, 0, st, A.rowptr.p, A.colidx.p, A.vals.p, A.n, x, y, S); break;
        case 16: hipLaunchKernelGGL(k_spmv_guard<16>, g, b, 0, st, A.rowptr.p, A.colidx.p, A.vals.p, A.n, x, y, S); break;
        case 32: hipLaunchKernelGGL(k_spmv_guard<32>, g, b, 0, st, A.rowptr.p, A.colidx.p, A.vals.p, A.n, x, y, S); break;
        default: hipLaunchKernelGGL(k_spmv_guard<64>, g, b, 0, st, A.rowptr.p, A.colidx.p, A.vals.p, A.n, x, y, S); break;
    }
}

// ConjugateGradientGaussPreSolver.SolveWithInitialGuess (CGGP/ConjugateGradientGaussPreSolver.cs:78-166)
void pcg_device(const Matrix &A, const double *d_b, const double *d_x0, double tol, int max_it, int use_jacobi,
                double *d_x, hipStream_t st, mg_pcg_result &res) {
    if (!std::isfinite(tol) || tol < 0.0) throw Error(MG_E_RANGE, "RelativeTolerance is out of range.");
    if (max_it < 0) throw Error(MG_E_RANGE, "MaxIterations: the maximum iteration count cannot be negative.");
    const int n = A.n;
    res = mg_pcg_result{0, 1, 0.0, 0.0};
    if (n == 0) return;
    if (d_x0) MG_HIP(hipMemcpyAsync(d_x, d_x0, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    else MG_HIP(hipMemsetAsync(d_x, 0, sizeof(double) * n, st));
    DevArray<double> r(n), ap(n), part(kRedBlocks), part2(kRedBlocks);
    DevArray<PcgScalars> S(1);
    MG_HIP(hipMemsetAsync(S.p, 0, sizeof(PcgScalars), st));
    const dim3 gv(nblk(n, kBS)), bv(kBS), gr(kRedBlocks);
    // r0 = b - A x0 (:115-117), ||r0||, ||b||
    launch_spmv(A, d_x, ap.p, st);
    hipLaunchKernelGGL(k_residual0, gv, bv, 0, st, d_b, ap.p, n, r.p);
    hipLaunchKernelGGL(k_dot_part, gr, bv, 0, st, r.p, r.p, n, part.p, (const PcgScalars *)nullptr);
    hipLaunchKernelGGL(k_dot_part, gr, bv, 0, st, d_b, d_b, n, part2.p, (const PcgScalars *)nullptr);
    hipLaunchKernelGGL(k_pcg_init, dim3(1), bv, 0, st, part.p, part2.p, tol, S.p);
    PcgScalars h{};
    MG_HIP(hipMemcpyAsync(&h, S.p, sizeof(h), hipMemcpyDeviceToHost, st));
    MG_HIP(hipStreamSynchronize(st));
    if (!h.done) {
        if (use_jacobi && A.diag_status)
            throw Error(MG_E_NOT_PD, "The diagonal element is zero or invalid while building the Jacobi preconditioner.");
        // single-reduction iteration: the start pass (kernel 0) forms w_0 = A M r_0 and the
        // partials of gamma_0 and delta_0; kernel it runs iteration it - 1 of the recurrences and
        // tests iteration it - 1's residual.  The iterate vectors alternate between two halves.
        DevArray<double> r2(n), w(2 * (size_t)n), s2(2 * (size_t)n), p(n);
        MG_HIP(hipMemsetAsync(w.p, 0, sizeof(double) * n, st));   // w_0 slot: zero before the start pass
        MG_HIP(hipMemsetAsync(s2.p, 0, sizeof(double) * n, st));  // s_{-1} = 0
        MG_HIP(hipMemsetAsync(p.p, 0, sizeof(double) * n, st));   // p_{-1} = 0
        const double *rb[2] = {r.p, r2.p}, *wb[2] = {w.p, w.p + n}, *sb[2] = {s2.p, s2.p + n};
        const int ng = pcg_dir_blocks(A);
        DevArray<double> cpart(2 * 3 * (size_t)ng);
        const int maxit = max_it > 0 ? max_it : n;
        const int check_every = 16;
        ScalarPoll &poll = ScalarPoll::get();
        auto launch = [&](int k) {
            const int a = k & 1, b1 = a ^ 1;
            const Cg1Vec v{rb[a], wb[a], sb[a], const_cast<double *>(rb[b1]), const_cast<double *>(wb[b1]),
                           const_cast<double *>(sb[b1])};
            const double *pin = cpart.p + 3 * (size_t)ng * (b1);
            double *pout = cpart.p + 3 * (size_t)ng * a;
            const dim3 g(ng), b(kBS);
            if (A.sell) {
                hipLaunchKernelGGL(k_pcg_cg1_sell, g, b, 0, st, A.sell_ptr.p, A.sell_ci.p, A.sell_vals.p, n, A.nslices,
                                   A.invdiag.p, use_jacobi, d_x, p.p, v, pin, ng, pout, k, S.p);
                return;
            }
            switch (A.lanes) {
                case 4: hipLaunchKernelGGL(k_pcg_cg1_csr<4>, g, b, 0, st, A.rowptr.p, A.colidx.p, A.vals.p, n, A.invdiag.p, use_jacobi, d_x, p.p, v, pin, ng, pout, k, S.p); break;
                case 8: hipLaunchKernelGGL(k_pcg_cg1_csr<8>, g, b, 0, st, A.rowptr.p, A.colidx.p, A.vals.p, n, A.invdiag.p, use_jacobi, d_x, p.p, v, pin, ng, pout, k, S.p); break;
                case 16: hipLaunchKernelGGL(k_pcg_cg1_csr<16>, g, b, 0, st, A.rowptr.p, A.colidx.p, A.vals.p, n, A.invdiag.p, use_jacobi, d_x, p.p, v, pin, ng, pout, k, S.p); break;
                case 32: hipLaunchKernelGGL(k_pcg_cg1_csr<32>, g, b, 0, st, A.rowptr.p, A.colidx.p, A.vals.p, n, A.invdiag.p, use_jacobi, d_x, p.p, v, pin, ng, pout, k, S.p); break;
                default: hipLaunchKernelGGL(k_pcg_cg1_csr<64>, g, b, 0, st, A.rowptr.p, A.colidx.p, A.vals.p, n, A.invdiag.p, use_jacobi, d_x, p.p, v, pin, ng, pout, k, S.p); break;
            }
        };
        launch(0);
        int it = 1;
        for (; it <= maxit; ++it) {
            launch(it);
            // iterations past convergence leave the state as it is (their kernels return at once)
            if (it % check_every == 0 && poll.post(S.p, it / check_every, st)) break;
        }
        if (it > maxit)
            hipLaunchKernelGGL(k_pcg_cg1_last, dim3(1), bv, 0, st, cpart.p + 3 * (size_t)ng * (maxit & 1), ng, maxit,
                               S.p);
        MG_HIP(hipGetLastError());
        MG_HIP(hipMemcpyAsync(&h, S.p, sizeof(h), hipMemcpyDeviceToHost, st));
        MG_HIP(hipStreamSynchronize(st));
        if (h.fail)
            throw Error(MG_E_NOT_PD, "p^T A p <= 0. The matrix is not positive definite, so the CG method cannot be applied.");
        if (!h.done) h.iters = maxit;
    }
    res.iterations = h.iters;
    res.residual_norm = h.rnorm;
    res.relative_residual = h.rnorm / h.bnorm;
    res.converged = res.relative_residual <= tol;
}

}  // namespace mg

// ---- flexible PCG with the learned preconditioner --------------------------------
//
// ConjugateGradientGaussPreSolver's iteration (CGGP/ConjugateGradientGaussPreSolver.cs:78-166)
// with z = G'G r, G the network-predicted factor (mpnn.hip).  The new direction is
// A-orthogonalised against the previous one, p = z + beta p with beta = -(z'Ap) / (p'Ap)
// (flexible CG; the same as PCG for a fixed preconditioner, and robust when the
// preconditioner changes at the fallback).  The stopping rule, the initial residual and the
// p'Ap <= 0 failure are the reference's.  Fallback to plain Jacobi-PCG (restart with
// p = z = D^{-1} r) when r'z <= 0 or when ||r|| has not reached a new minimum for `stall`
// iterations.

namespace mg {

// convergence test (:151-153) and stall tracking
__global__ void k_fcg_norm(const double *__restrict__ part, int it, int stall, PcgScalars *__restrict__ S) {
    __shared__ double sm[kBS];
    if (S->done) return;
    const double rr = finish_sum(part, sm);
    if (threadIdx.x == 0) {
        S->rnorm = sqrt(rr);
        S->iters = it;
        if (S->rnorm <= S->thr) {
            S->done = 1;
        } else if (!S->fallback) {
            if (S->rnorm < S->best) {
                S->best = S->rnorm;
                S->best_it = it;
            } else if (it - S->best_it >= stall) {
                S->fallback = 1;
                S->restart = 1;
            }
        }
    }
}

// the same test over the nparts partials of the two-kernel step
__global__ void k_fcg_norm_n(const double *__restrict__ part, int nparts, int it, int stall,
                             PcgScalars *__restrict__ S) {
    __shared__ double sm4[4];
    if (S->done) return;
    const double rr = sum_partials(part, nparts, sm4);
    if (threadIdx.x == 0) {
        S->rnorm = sqrt(rr);
        S->iters = it;
        if (S->rnorm <= S->thr) {
            S->done = 1;
        } else if (!S->fallback) {
            if (S->rnorm < S->best) {
                S->best = S->rnorm;
                S->best_it = it;
            } else if (it - S->best_it >= stall) {
                S->fallback = 1;
                S->restart = 1;
            }
        }
    }
}

// Jacobi z = D^{-1} r with partials of r'z and z'Ap; active only after the fallback
__global__ __launch_bounds__(256) void k_fcg_jacobi(const double *__restrict__ invd, const double *__restrict__ r,
                                                    const double *__restrict__ ap, double *__restrict__ z, int n,
                                                    double *__restrict__ rz_part, double *__restrict__ zap_part,
                                                    const PcgScalars *__restrict__ S) {
    __shared__ double sm[kBS];
    if (S->done || !S->fallback) return;
    double a = 0.0, b = 0.0;
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
        const double zi = invd[i] * r[i];
        z[i] = zi;
        a += r[i] * zi;
        if (ap) b += zi * ap[i];
    }
    const double ta = block_sum(a, sm);
    __syncthreads();
    const double tb = block_sum(b, sm);
    if (threadIdx.x == 0) {
        rz_part[blockIdx.x] = ta;
        zap_part[blockIdx.x] = tb;
    }
}

// first direction: rz = r'z; a non-positive r'z switches to Jacobi before the first step
__global__ void k_fcg_start(const double *__restrict__ rz_part, int nparts, PcgScalars *__restrict__ S, int second) {
    __shared__ double sm[kBS];
    if (S->done) return;
    if (second && !S->fallback) return;
    const double rz = finish_sum_n(rz_part, nparts, sm);
    if (threadIdx.x == 0) {
        S->rz = rz;
        S->best = S->rnorm;
        S->best_it = 0;
        if (!second && !(rz > 0.0)) S->fallback = 1;
    }
}

// beta = -(z'Ap) / (p'Ap) (0 on a restart); rz = r'z for the next step length
__global__ void k_fcg_beta(const double *__restrict__ rz_part, const double *__restrict__ zap_part, int nparts,
                           PcgScalars *__restrict__ S) {
    __shared__ double sm4[4];
    if (S->done) return;
    const double rz = sum_partials(rz_part, nparts, sm4);
    const double zap = sum_partials(zap_part, nparts, sm4);
    if (threadIdx.x == 0) {
        S->rz = rz;
        if (S->restart || !(rz > 0.0)) {
            S->beta = 0.0;
            S->restart = 0;
            if (!(rz > 0.0) && !S->fallback) {
                S->fallback = 1;
                S->restart = 1;
            }
        } else {
            S->beta = -zap / S->pap;
        }
    }
}

// slot 0 of the fused flexible PCG's state from the start (r'z, ||r0||, a fallback at the start)
__global__ void k_fcg_state0(const PcgScalars *__restrict__ S, FcgState *__restrict__ fs) {
    fs[0].rnorm = S->rnorm;
    fs[0].best = S->best;
    fs[0].best_it = 0;
    fs[0].rz2 = S->rz;
    fs[0].beta = 0.0;
    fs[0].fallback1 = fs[0].fallback2 = S->fallback;
    fs[0].restart1 = fs[0].restart2 = 0;
}

void fcg_learned_device(Matrix &A, Mpnn &net, const double *d_b, const double *d_x0, double tol, int max_it,
                        int stall_window, double *d_x, hipStream_t st, mg_pcg_result &res, int &fell_back) {
    if (!std::isfinite(tol) || tol < 0.0) throw Error(MG_E_RANGE, "RelativeTolerance is out of range.");
    if (max_it < 0) throw Error(MG_E_RANGE, "MaxIterations: the maximum iteration count cannot be negative.");
    const int n = A.n;
    res = mg_pcg_result{0, 1, 0.0, 0.0};
    fell_back = 0;
    if (n == 0) return;
    if (A.diag_status)
        throw Error(MG_E_NOT_PD, "The diagonal element is zero or invalid while building the Jacobi preconditioner.");
    mp_graph_prepare(A, st);
    if (d_x0) MG_HIP(hipMemcpyAsync(d_x, d_x0, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    else MG_HIP(hipMemsetAsync(d_x, 0, sizeof(double) * n, st));
    // the preconditioner kernels write one partial sum per block of their full grid
    const int npart = mpnn_partial_blocks(n);
    DevArray<double> r(n), z(n), p(n), p2(n), ap(n), part(std::max(npart, kRedBlocks)),
        part2(std::max(npart, kRedBlocks));
    DevArray<PcgScalars> S(1);
    MG_HIP(hipMemsetAsync(S.p, 0, sizeof(PcgScalars), st));
    const dim3 gv(nblk(n, kBS)), bv(kBS), gr(kRedBlocks), gp(npart);
    launch_spmv(A, d_x, ap.p, st);
    hipLaunchKernelGGL(k_residual0, gv, bv, 0, st, d_b, ap.p, n, r.p);
    hipLaunchKernelGGL(k_dot_part, gr, bv, 0, st, r.p, r.p, n, part.p, (const PcgScalars *)nullptr);
    hipLaunchKernelGGL(k_dot_part, gr, bv, 0, st, d_b, d_b, n, part2.p, (const PcgScalars *)nullptr);
    hipLaunchKernelGGL(k_pcg_init, dim3(1), bv, 0, st, part.p, part2.p, tol, S.p);
    PcgScalars h{};
    MG_HIP(hipMemcpyAsync(&h, S.p, sizeof(h), hipMemcpyDeviceToHost, st));
    MG_HIP(hipStreamSynchronize(st));
    if (!h.done) {
        // factor values from the BF16 network (once per matrix and network), z0 = G'G r0
        mpnn_factor_bf16(net, A, st);
        mpnn_apply_factor(net, A, r.p, z.p, S.p, nullptr, part.p, part2.p, st);
        hipLaunchKernelGGL(k_fcg_start, dim3(1), bv, 0, st, part.p, npart, S.p, 0);
        // on a fallback at the start, the same call now computes the Jacobi z0
        hipLaunchKernelGGL(k_fcg_jacobi, gp, bv, 0, st, A.invdiag.p, r.p, (const double *)nullptr, z.p, n, part.p,
                           part2.p, S.p);
        hipLaunchKernelGGL(k_fcg_start, dim3(1), bv, 0, st, part.p, npart, S.p, 1);
        MG_HIP(hipMemcpyAsync(p.p, z.p, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
        const int maxit = max_it > 0 ? max_it : n;
        const int check_every = 16;
        const int stall = stall_window > 0 ? stall_window : 50;
        ScalarPoll &poll = ScalarPoll::get();
        bool fused = false;
        if (A.sell && A.mp && A.mp->nslices) {
            // four kernels per iteration: the direction (p = z + beta p in the gather, beta from
            // the previous iteration's partials; beta = 0 and p_old = 0 at the start) with the
            // SpMV and p'Ap partials; alpha and the x / r update with ||r||^2 partials; y = G r
            // and z = G' y (or Jacobi), each with the residual-norm test of the step folded in
            MG_HIP(hipMemsetAsync(p2.p, 0, sizeof(double) * n, st));
            const int ndir = pcg_dir_blocks(A), nstep = pcg_step_blocks(n), nfcg = mpnn_fcg_partials(A);
            DevArray<double> pdir(ndir), prr(nstep);
            DevArray<FcgState> fs(2);
            hipLaunchKernelGGL(k_fcg_state0, dim3(1), dim3(1), 0, st, S.p, fs.p);
            double *p_old = p2.p, *p_new = p.p;
            fused = true;
            for (int it = 1; it <= maxit; ++it) {
                FcgState *prev = fs.p + ((it - 1) & 1), *cur = fs.p + (it & 1);
                hipLaunchKernelGGL(k_pcg_dir_sell, dim3(ndir), bv, 0, st, A.sell_ptr.p, A.sell_ci.p, A.sell_vals.p, n,
                                   A.nslices, z.p, p_old, p_new, ap.p, part.p, part2.p, nfcg, it, pdir.p, S.p, prev);
                hipLaunchKernelGGL(k_pcg_step, dim3(nstep), bv, 0, st, d_x, r.p, p_new, ap.p, n, pdir.p, ndir, prr.p,
                                   S.p, (const double *)&prev->rz2);
                if (!mpnn_fcg_apply(net, A, r.p, z.p, ap.p, prr.p, nstep, it, stall, prev, cur, S.p, part.p, part2.p,
                                    st))
                    throw Error(MG_E_DEVICE, "The learned factor has no sliced copy for the fused iteration.");
                std::swap(p_old, p_new);
                if (it % check_every == 0 && poll.post(S.p, it / check_every, st)) break;
            }
        }
        for (int it = 1; !fused && it <= maxit; ++it) {
            launch_spmv_guard(A, p.p, ap.p, S.p, st);
            hipLaunchKernelGGL(k_dot_part, gr, bv, 0, st, p.p, ap.p, n, part.p, S.p);
            hipLaunchKernelGGL(k_pcg_alpha, dim3(1), bv, 0, st, part.p, S.p);
            hipLaunchKernelGGL(k_pcg_update, gr, bv, 0, st, d_x, r.p, p.p, ap.p, n, part.p, S.p);
            hipLaunchKernelGGL(k_fcg_norm, dim3(1), bv, 0, st, part.p, it, stall, S.p);
            mpnn_apply_factor(net, A, r.p, z.p, S.p, ap.p, part.p, part2.p, st);  // Jacobi after a fallback
            hipLaunchKernelGGL(k_fcg_beta, dim3(1), bv, 0, st, part.p, part2.p, npart, S.p);
            hipLaunchKernelGGL(k_pcg_pupdate, gv, bv, 0, st, z.p, p.p, n, S.p);
            if (it % check_every == 0 || it == maxit) {
                MG_HIP(hipMemcpyAsync(&h, S.p, sizeof(h), hipMemcpyDeviceToHost, st));
                MG_HIP(hipStreamSynchronize(st));
                if (h.done) break;
            }
        }
        MG_HIP(hipGetLastError());
        MG_HIP(hipMemcpyAsync(&h, S.p, sizeof(h), hipMemcpyDeviceToHost, st));
        MG_HIP(hipStreamSynchronize(st));
        if (h.fail)
            throw Error(MG_E_NOT_PD, "p^T A p <= 0. The matrix is not positive definite, so the CG method cannot be applied.");
        if (!h.done) h.iters = maxit;
        fell_back = h.fallback;
    }
    res.iterations = h.iters;
    res.residual_norm = h.rnorm;
    res.relative_residual = h.rnorm / h.bnorm;
    res.converged = res.relative_residual <= tol;
}

}  // namespace mg
