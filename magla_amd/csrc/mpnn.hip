// mpnn.hip -- learned preconditioner: a small message-passing network on the matrix
// graph predicts a sparse lower-triangular factor G (pattern: columns j <= i within two
// hops of i), and z = G'G r corrects CGGP's Jacobi preconditioner
// (CGGP/JacobiPreconditioner.cs:12-56) inside a flexible PCG.  MAGLA has no counterpart;
// the network is defined here:
//
//   edges    e1_ij = a_ij / sqrt(a_ii a_jj) (j != i);  e2_ij = sum_k e1_ik e1_kj
//   nodes    x_i = [1, c_i, s_i], c_i = log(a_ii / mean a_ii), s_i = sum_j |e1_ij|
//   h0_i   = W_in x_i                                                  (width 32)
//   layer l: m_i = sum_j e1_ij h_j ;  h'_i = act(W1_l h_i + W2_l m_i)  (act = ReLU for
//            l < 2, identity for l = 2)
//   factor   G_ii = a_ii^{-1/2} exp(w_t . h3_i)
//            G_ij = a_jj^{-1/2} (e1_ij p1_ij + e2_ij p2_ij),
//            p_ij = u . (h3_i * h3_j) + v . h3_i + w . h3_j  with (u1, v1, w1) or (u2, v2, w2)
//
// Zero readouts give G = D^{-1/2}, i.e. Jacobi.  z = G'G r is linear, symmetric and positive
// definite for any parameters.  Training (FP32 SIMT kernels, Adam) minimises, over (r, e = A^{-1} r)
// pairs, the squared A-norm error of z relative to e, L = (z'Az - 2 z'r) / (r'e) + 1, plus
// kappa times the log Kaporin condition number of G A G', log(tr(G A G') / n) - (2/n) sum log G_ii,
// which keeps small eigenvalues of the preconditioned operator away from zero.  Inference runs
// the network once per matrix in BF16 (weights and activations) with FP32 accumulation on MFMA
// (v_mfma_f32_16x16x16_bf16: per 16-node tile, [h | m] (16 x 64) x [W1 | W2]^T (64 x 32)); the
// factor values are then stored in FP32 and the solves apply G and G' with FP64 accumulation
// and FP64 vectors.  Every reduction has a fixed grid and order.
#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

#include "device.h"
#include "reduce.h"

namespace mg {

namespace {

constexpr int W = kMpW;
constexpr int kNB = 256 / W;  // nodes per 256-thread block in the SIMT kernels

typedef short mp_s4 __attribute__((ext_vector_type(4)));
typedef float mp_f4 __attribute__((ext_vector_type(4)));

inline int nblk(int64_t t, int b) { return (int)((t + b - 1) / b); }

__device__ inline uint16_t f2bf(float x) {  // round to nearest even
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint32_t b) { return __uint_as_float(b << 16); }

__device__ inline bool gated(const PcgScalars *S) { return S && (S->done || S->fallback); }
__device__ inline float to_f32(float v) { return v; }
__device__ inline float to_f32(uint16_t v) { return bf2f(v); }

// ---- graph data ------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_sum_final(const double *__restrict__ part, int nparts, double scale,
                                                   double *__restrict__ out) {
    __shared__ double sm[256];
    double acc = 0.0;
    for (int b = threadIdx.x; b < nparts; b += 256) acc += part[b];
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sm[threadIdx.x] += sm[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sm[0] * scale;
}

// One message-passing layer: m_i = sum_j ew_ij h_j, pre = W1 h_i + W2 m_i, h' = act(pre).
// 32 lanes per node (lane = feature), 8 nodes per pass of a block; the layer's weights in LDS.
__global__ __launch_bounds__(256) void k_mp_layer(const float *__restrict__ wl, const int *__restrict__ rp,
                                                  const int *__restrict__ ci, const float *__restrict__ ew, int n,
                                                  const float *__restrict__ Hin, float *__restrict__ Mout,
                                                  float *__restrict__ Pre, float *__restrict__ Hout, int relu) {
    __shared__ float Ws[2][W][W + 1];
    __shared__ float hs[kNB][W], ms[kNB][W];
    for (int idx = threadIdx.x; idx < 2 * W * W; idx += 256) Ws[idx / (W * W)][(idx / W) % W][idx % W] = wl[idx];
    __syncthreads();
    const int g = threadIdx.x / W, f = threadIdx.x % W;
    for (int base = blockIdx.x * kNB; base < n; base += gridDim.x * kNB) {
        const int i = base + g;
        const bool valid = i < n;
        float h = 0.0f, m = 0.0f;
        if (valid) {
            h = Hin[(int64_t)i * W + f];
            for (int k = rp[i]; k < rp[i + 1]; ++k) m += ew[k] * Hin[(int64_t)ci[k] * W + f];
        }
        hs[g][f] = h;
        ms[g][f] = m;
        __syncthreads();
        float pre = 0.0f;
#pragma unroll 8
        for (int q = 0; q < W; ++q) pre += Ws[0][f][q] * hs[g][q] + Ws[1][f][q] * ms[g][q];
        if (valid) {
            const int64_t e = (int64_t)i * W + f;
            Mout[e] = m;
            Pre[e] = pre;
            Hout[e] = (relu && !(pre > 0.0f)) ? 0.0f : pre;
        }
        __syncthreads();
    }
}

// ---- loss and FP32 backward ----------------------------------------------------

__global__ __launch_bounds__(256) void k_mp_loss_part(const double *__restrict__ z, const double *__restrict__ az,
                                                      const double *__restrict__ r, const double *__restrict__ e,
                                                      int n, double *__restrict__ part) {
    __shared__ double sm[3][256];
    double a = 0.0, b = 0.0, c = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        a += z[i] * az[i];
        b += z[i] * r[i];
        c += r[i] * e[i];
    }
    sm[0][threadIdx.x] = a;
    sm[1][threadIdx.x] = b;
    sm[2][threadIdx.x] = c;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int q = 0; q < 3; ++q) sm[q][threadIdx.x] += sm[q][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x < 3) part[threadIdx.x * kRedBlocks + blockIdx.x] = sm[threadIdx.x][0];
}

// scal[1] = r'e; scal[2] += loss / ns
__global__ __launch_bounds__(256) void k_mp_loss_final(const double *__restrict__ part, int ns,
                                                       double *__restrict__ scal) {
    __shared__ double sm[3][256];
    double v[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < kRedBlocks; b += 256)
        for (int q = 0; q < 3; ++q) v[q] += part[q * kRedBlocks + b];
    for (int q = 0; q < 3; ++q) sm[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int q = 0; q < 3; ++q) sm[q][threadIdx.x] += sm[q][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        scal[1] = sm[2][0];
        scal[2] += ((sm[0][0] - 2.0 * sm[1][0]) / sm[2][0] + 1.0) / ns;
    }
}

// Fixed-order sum of the 8 node groups' per-feature values of a block, added into part.
__device__ inline void group_reduce_add(float v, float (*sm)[W], float *__restrict__ dst) {
    const int g = threadIdx.x / W, f = threadIdx.x % W;
    sm[g][f] = v;
    __syncthreads();
    if (g == 0) {
        float t = 0.0f;
        for (int q = 0; q < kNB; ++q) t += sm[q][f];
        dst[f] += t;
    }
    __syncthreads();
}

// Backward of one layer: dpre = dh (x ReLU mask); dh_self = W1' dpre, dm = W2' dpre;
// weight gradients dW1 += dpre h', dW2 += dpre m' (per block, into part).
__global__ __launch_bounds__(256) void k_mp_layer_bwd(const float *__restrict__ wl, int n, const float *__restrict__ dH,
                                                      const float *__restrict__ Pre, const float *__restrict__ Hin,
                                                      const float *__restrict__ Mb, int relu, float *__restrict__ dHp,
                                                      float *__restrict__ dM, float *__restrict__ part) {
    __shared__ float Ws[2][W][W + 1];
    __shared__ float dps[kNB][W], hs[kNB][W], ms[kNB][W];
    __shared__ float red[kNB][W][W + 1];
    for (int idx = threadIdx.x; idx < 2 * W * W; idx += 256) Ws[idx / (W * W)][(idx / W) % W][idx % W] = wl[idx];
    __syncthreads();
    const int g = threadIdx.x / W, f = threadIdx.x % W;
    float a1[W], a2[W];
#pragma unroll
    for (int q = 0; q < W; ++q) a1[q] = a2[q] = 0.0f;
    for (int base = blockIdx.x * kNB; base < n; base += gridDim.x * kNB) {
        const int i = base + g;
        const bool valid = i < n;
        const int64_t e = (int64_t)i * W + f;
        float dp = 0.0f, h = 0.0f, m = 0.0f;
        if (valid) {
            dp = (relu && !(Pre[e] > 0.0f)) ? 0.0f : dH[e];
            h = Hin[e];
            m = Mb[e];
        }
        dps[g][f] = dp;
        hs[g][f] = h;
        ms[g][f] = m;
        __syncthreads();
        float s1 = 0.0f, s2 = 0.0f;
#pragma unroll 8
        for (int q = 0; q < W; ++q) {
            s1 += Ws[0][q][f] * dps[g][q];
            s2 += Ws[1][q][f] * dps[g][q];
        }
        if (valid) {
            dHp[e] = s1;
            dM[e] = s2;
        }
#pragma unroll
        for (int q = 0; q < W; ++q) {
            a1[q] += dp * hs[g][q];
            a2[q] += dp * ms[g][q];
        }
        __syncthreads();
    }
    float *dst = part + (int64_t)blockIdx.x * kMpP;
    for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
        for (int q = 0; q < W; ++q) red[g][f][q] = pass == 0 ? a1[q] : a2[q];
        __syncthreads();
        for (int idx = threadIdx.x; idx < W * W; idx += 256) {
            const int ff = idx / W, q = idx % W;
            float t = 0.0f;
            for (int gg = 0; gg < kNB; ++gg) t += red[gg][ff][q];
            dst[pass * W * W + idx] += t;
        }
        __syncthreads();
    }
}

// transpose of the aggregation (ew symmetric): dh_i += sum_j ew_ij dm_j
__global__ __launch_bounds__(256) void k_mp_agg_bwd(const int *__restrict__ rp, const int *__restrict__ ci,
                                                    const float *__restrict__ ew, int n, const float *__restrict__ dM,
                                                    float *__restrict__ dHp) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int i = (int)(e / W), f = (int)(e % W);
    if (i >= n) return;
    float acc = dHp[e];
    for (int k = rp[i]; k < rp[i + 1]; ++k) acc += ew[k] * dM[(int64_t)ci[k] * W + f];
    dHp[e] = acc;
}

__global__ void k_mp_grad_reduce(const float *__restrict__ part, int nparts, float *__restrict__ grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMpP) return;
    float t = 0.0f;
    for (int b = 0; b < nparts; ++b) t += part[(int64_t)b * kMpP + p];
    grad[p] = t;
}

__global__ void k_fill_f32(float *__restrict__ p, int64_t count, float v) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < count) p[e] = v;
}
__global__ void k_fill_f64(double *__restrict__ p, int count, double v) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < count) p[e] = v;
}

__global__ void k_mp_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                          float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMpP) return;
    const float gp = g[p];
    const float mp = b1 * m[p] + (1.0f - b1) * gp;
    const float vp = b2 * v[p] + (1.0f - b2) * gp * gp;
    m[p] = mp;
    v[p] = vp;
    w[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);
}

// ---- BF16 inference --------------------------------------------------------------

// One layer on MFMA.  Each wave takes 16-node tiles: lane l gathers features 8(l%4)..+7
// of node l/4 (own h and the weighted neighbour sum m, FP32), writes the bf16 row
// [h | m] to LDS, then 2 x 4 v_mfma_f32_16x16x16_bf16 give the 16 x 32 outputs.
__global__ __launch_bounds__(256) void k_mpb_layer(const uint16_t *__restrict__ Bw, const int *__restrict__ rp,
                                                   const int *__restrict__ ci, const float *__restrict__ ew, int n,
                                                   const uint16_t *__restrict__ Hin, uint16_t *__restrict__ Hout,
                                                   int relu, const PcgScalars *__restrict__ S) {
    if (gated(S)) return;
    __shared__ __align__(16) uint16_t As[4][16][72];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    mp_s4 b[2][4];
#pragma unroll
    for (int nb = 0; nb < 2; ++nb)
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) b[nb][kk] = reinterpret_cast<const mp_s4 *>(Bw)[(nb * 4 + kk) * 64 + lane];
    const int ntiles = (n + 15) / 16;
    const int row = lane >> 2, c = lane & 3;
    for (int t0 = blockIdx.x * 4; t0 < ntiles; t0 += gridDim.x * 4) {
        const int tile = t0 + wave;
        const int node = tile * 16 + row;
        uint4 hv = make_uint4(0, 0, 0, 0);
        float m[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) m[q] = 0.0f;
        if (tile < ntiles && node < n) {
            hv = *reinterpret_cast<const uint4 *>(Hin + (int64_t)node * W + 8 * c);
            for (int k = rp[node]; k < rp[node + 1]; ++k) {
                const float wgt = ew[k];
                const uint4 hj = *reinterpret_cast<const uint4 *>(Hin + (int64_t)ci[k] * W + 8 * c);
                const uint32_t v4[4] = {hj.x, hj.y, hj.z, hj.w};
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    m[2 * q] += wgt * bf2f(v4[q] & 0xFFFFu);
                    m[2 * q + 1] += wgt * bf2f(v4[q] >> 16);
                }
            }
        }
        uint4 mv;
        mv.x = (uint32_t)f2bf(m[0]) | ((uint32_t)f2bf(m[1]) << 16);
        mv.y = (uint32_t)f2bf(m[2]) | ((uint32_t)f2bf(m[3]) << 16);
        mv.z = (uint32_t)f2bf(m[4]) | ((uint32_t)f2bf(m[5]) << 16);
        mv.w = (uint32_t)f2bf(m[6]) | ((uint32_t)f2bf(m[7]) << 16);
        *reinterpret_cast<uint4 *>(&As[wave][row][8 * c]) = hv;
        *reinterpret_cast<uint4 *>(&As[wave][row][W + 8 * c]) = mv;
        __syncthreads();
        mp_f4 acc0 = {0.0f, 0.0f, 0.0f, 0.0f}, acc1 = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const mp_s4 a = *reinterpret_cast<const mp_s4 *>(&As[wave][lane & 15][16 * kk + 4 * (lane >> 4)]);
            acc0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[0][kk], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[1][kk], acc1, 0, 0, 0);
        }
        if (tile < ntiles) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int nd = tile * 16 + 4 * (lane >> 4) + i;
                if (nd < n) {
                    float v0 = acc0[i], v1 = acc1[i];
                    if (relu) {
                        v0 = v0 > 0.0f ? v0 : 0.0f;
                        v1 = v1 > 0.0f ? v1 : 0.0f;
                    }
                    Hout[(int64_t)nd * W + (lane & 15)] = f2bf(v0);
                    Hout[(int64_t)nd * W + 16 + (lane & 15)] = f2bf(v1);
                }
            }
        }
        __syncthreads();
    }
}

// ---- learned-factor kernels -------------------------------------------------------

__global__ __launch_bounds__(256) void k_mp_input(const float *__restrict__ x, const float *__restrict__ w, int n,
                                                  float *__restrict__ H0) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int i = (int)(e / W), f = (int)(e % W);
    if (i >= n) return;
    H0[e] = w[3 * f] * x[3 * i] + w[3 * f + 1] * x[3 * i + 1] + w[3 * f + 2] * x[3 * i + 2];
}

// readout values of one factor entry from the node states (32 lanes = features, reduced over
// the lanes of the node's group)
template <typename H, typename R>
__device__ inline float factor_entry(int i, int j, const H *hi, const H *hj, const R *rd, int f, float e1, float e2,
                                     const double *__restrict__ diag, double &out) {
    (void)i;
    const float a = to_f32(hi[f]), b = to_f32(hj[f]);
    float p1 = to_f32(rd[W + f]) * a * b + to_f32(rd[2 * W + f]) * a + to_f32(rd[3 * W + f]) * b;
    float p2 = to_f32(rd[4 * W + f]) * a * b + to_f32(rd[5 * W + f]) * a + to_f32(rd[6 * W + f]) * b;
    for (int o = W / 2; o > 0; o >>= 1) {
        p1 += __shfl_xor(p1, o, W);
        p2 += __shfl_xor(p2, o, W);
    }
    out = ((double)e1 * p1 + (double)e2 * p2) / sqrt(diag[j]);
    return 0.0f;
}

// G values of every row from h3 (FP32 or BF16 states and readout weights); 32 lanes per node.
template <typename H, typename R, typename G>
__global__ __launch_bounds__(256) void k_mp_readout(const H *__restrict__ h3, const R *__restrict__ rd,
                                                    const int *__restrict__ gptr, const int *__restrict__ gcol,
                                                    const float *__restrict__ ge1, const float *__restrict__ ge2,
                                                    const double *__restrict__ diag, int n, G *__restrict__ gv) {
    const int g = threadIdx.x / W, f = threadIdx.x % W;
    for (int base = blockIdx.x * kNB; base < n; base += gridDim.x * kNB) {
        const int i = base + g;
        if (i >= n) continue;  // whole 32-lane group
        const H *hi = h3 + (int64_t)i * W;
        for (int e = gptr[i]; e < gptr[i + 1]; ++e) {
            const int j = gcol[e];
            double val;
            if (j == i) {
                float t = to_f32(rd[f]) * to_f32(hi[f]);
                for (int o = W / 2; o > 0; o >>= 1) t += __shfl_xor(t, o, W);
                val = exp((double)t) / sqrt(diag[i]);
            } else {
                factor_entry(i, j, hi, h3 + (int64_t)j * W, rd, f, ge1[e], ge2[e], diag, val);
            }
            if (f == 0) gv[e] = (G)val;
        }
    }
}

// Row products of the factor: kRowLanes lanes per row (row entries strided over the lanes,
// fixed xor-tree per row), so a wave reads consecutive entries of a few rows at once.
constexpr int kRowLanes = 8;

// y_i = sum_e G_e v_j over row i
template <typename G>
__global__ __launch_bounds__(256) void k_g_apply(const int *__restrict__ gptr, const int *__restrict__ gcol,
                                                 const G *__restrict__ gv, const double *__restrict__ v, int n,
                                                 double *__restrict__ y, const PcgScalars *__restrict__ S) {
    if (gated(S)) return;
    const int gid = blockIdx.x * 256 + threadIdx.x;
    const int i = gid / kRowLanes, sub = gid % kRowLanes;
    double acc = 0.0;
    if (i < n) {
        // the next entry's column and value load with the current entry's v (indices clamped
        // to an entry of the pattern, which holds at least one)
        const int ee = gptr[i + 1], elast = max(ee - 1, 0);
        int e = gptr[i] + sub;
        int col = gcol[min(e, elast)];
        G gval = gv[min(e, elast)];
        for (; e < ee; e += kRowLanes) {
            const int cn = gcol[min(e + kRowLanes, elast)];
            const G gn = gv[min(e + kRowLanes, elast)];
            acc += (double)gval * v[col];
            col = cn;
            gval = gn;
        }
    }
    for (int o = kRowLanes / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kRowLanes);
    if (i < n && sub == 0) y[i] = acc;
}

// z_j = sum over column j of G of G_ij y_i (transposed pattern; gvt = G values in that order,
// or gv through gtidx when gvt is null), with optional per-block partial sums of r.z and z.ap
// Rows are strided over the grid (any grid size), so the partial sums have one value per
// block of a fixed grid.  After a fallback (S->fallback) it computes the Jacobi z = D^{-1} r.
template <typename G>
__global__ __launch_bounds__(256) void k_gt_apply(const int *__restrict__ gtptr, const int *__restrict__ gtrow,
                                                  const int *__restrict__ gtidx, const G *__restrict__ gv,
                                                  const G *__restrict__ gvt, const double *__restrict__ y, int n,
                                                  double *__restrict__ z, const double *__restrict__ r,
                                                  const double *__restrict__ ap, double *__restrict__ rz_part,
                                                  double *__restrict__ zap_part, const double *__restrict__ invd,
                                                  const PcgScalars *__restrict__ S) {
    if (S && S->done) return;
    const bool jacobi = S && S->fallback;
    __shared__ double sm[2][256];
    const int sub = threadIdx.x % kRowLanes;
    double arz = 0.0, azap = 0.0;
    for (int jb = (blockIdx.x * 256 + threadIdx.x) / kRowLanes; jb < n + 0; jb += gridDim.x * (256 / kRowLanes)) {
        const int j = jb;
        double acc = 0.0;
        if (jacobi) {
            acc = sub == 0 ? invd[j] * r[j] : 0.0;
        } else if (gvt) {
            // the next entry's row and value load with the current entry's y (indices clamped
            // to an entry of the pattern, which holds at least one)
            const int te = gtptr[j + 1], tlast = max(te - 1, 0);
            int t = gtptr[j] + sub;
            int row = gtrow[min(t, tlast)];
            G gval = gvt[min(t, tlast)];
            for (; t < te; t += kRowLanes) {
                const int rn = gtrow[min(t + kRowLanes, tlast)];
                const G gn = gvt[min(t + kRowLanes, tlast)];
                acc += (double)gval * y[row];
                row = rn;
                gval = gn;
            }
        } else {
            for (int t = gtptr[j] + sub; t < gtptr[j + 1]; t += kRowLanes) acc += (double)gv[gtidx[t]] * y[gtrow[t]];
        }
        for (int o = kRowLanes / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kRowLanes);
        if (sub == 0) {
            z[j] = acc;
            if (rz_part) {
                arz += r[j] * acc;
                if (ap) azap += acc * ap[j];
            }
        }
    }
    if (!rz_part) return;
    sm[0][threadIdx.x] = arz;
    sm[1][threadIdx.x] = azap;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            sm[0][threadIdx.x] += sm[0][threadIdx.x + h];
            sm[1][threadIdx.x] += sm[1][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        rz_part[blockIdx.x] = sm[0][0];
        zap_part[blockIdx.x] = sm[1][0];
    }
}

__global__ void k_gather_sell(const int *__restrict__ src, int64_t count, const float *__restrict__ gv,
                              float *__restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < count) out[e] = src[e] >= 0 ? gv[src[e]] : 0.0f;
}

// Packed sliced entries for the fused PCG: the value rounded to BF16 (nearest, ties to even) in
// the high 16 bits, the index difference from the slot's row in the low 16 bits.
__global__ void k_pack_sell(const int *__restrict__ src, const uint16_t *__restrict__ dl, int64_t count,
                            const float *__restrict__ gv, uint32_t *__restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    const uint32_t u = src[e] >= 0 ? __float_as_uint(gv[src[e]]) : 0u;
    const uint32_t hi = (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;
    out[e] = hi | (uint32_t)dl[e];
}

// y = G v on the sliced ELL copy of G: wave = 64-row slice, lane = row, the row's entries
// accumulated in order in float64
__global__ __launch_bounds__(256) void k_g_apply_sell(const int *__restrict__ sp, const int *__restrict__ scol,
                                                      const float *__restrict__ sv, int n, int nslices,
                                                      const double *__restrict__ v, double *__restrict__ y,
                                                      const PcgScalars *__restrict__ S) {
    if (gated(S)) return;
    const int lane = threadIdx.x & 63, slice = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (slice >= nslices) return;
    const int base = sp[slice], w = (sp[slice + 1] - base) >> 6, row = 64 * slice + lane;
    double acc = 0.0;
    for (int k = 0; k < w; ++k) {
        const int e = base + 64 * k + lane;
        acc += (double)sv[e] * v[scol[e]];
    }
    if (row < n) y[row] = acc;
}

// z = G' y on the sliced ELL copy of G' (Jacobi z = D^{-1} r after a fallback), with block
// partials of r'z and z'Ap over a fixed grid (slices strided over the blocks)
__global__ __launch_bounds__(256) void k_gt_apply_sell(const int *__restrict__ sp, const int *__restrict__ srow,
                                                       const float *__restrict__ sv, int n, int nslices,
                                                       const double *__restrict__ y, double *__restrict__ z,
                                                       const double *__restrict__ r, const double *__restrict__ ap,
                                                       double *__restrict__ rz_part, double *__restrict__ zap_part,
                                                       const double *__restrict__ invd,
                                                       const PcgScalars *__restrict__ S) {
    if (S && S->done) return;
    __shared__ double sm[2][4];
    const bool jacobi = S && S->fallback;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double arz = 0.0, azap = 0.0;
    for (int slice = blockIdx.x * 4 + wave; slice < nslices; slice += gridDim.x * 4) {
        const int j = 64 * slice + lane;
        double acc = 0.0;
        if (jacobi) {
            if (j < n) acc = invd[j] * r[j];
        } else {
            const int base = sp[slice], w = (sp[slice + 1] - base) >> 6;
            for (int k = 0; k < w; ++k) {
                const int e = base + 64 * k + lane;
                acc += (double)sv[e] * y[srow[e]];
            }
        }
        if (j < n) {
            z[j] = acc;
            if (rz_part) {
                arz += r[j] * acc;
                if (ap) azap += acc * ap[j];
            }
        }
    }
    if (!rz_part) return;
    for (int o = 32; o > 0; o >>= 1) {
        arz += __shfl_xor(arz, o, 64);
        azap += __shfl_xor(azap, o, 64);
    }
    if (lane == 0) {
        sm[0][wave] = arz;
        sm[1][wave] = azap;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        rz_part[blockIdx.x] = (sm[0][0] + sm[0][1]) + (sm[0][2] + sm[0][3]);
        zap_part[blockIdx.x] = (sm[1][0] + sm[1][1]) + (sm[1][2] + sm[1][3]);
    }
}

// sum_k sv[e_k] * x[idx[e_k]] over one sliced-ELL row (e_k = base + 64 k + lane, k < w), in
// batches of eight entries whose index, value and gathered loads are issued together; the
// products are added in k order
__device__ __forceinline__ double sell_row_dot(const int *__restrict__ idx, const float *__restrict__ sv, int base,
                                               int w, int lane, const double *__restrict__ x) {
    double acc = 0.0;
    for (int k0 = 0; k0 < w; k0 += 8) {
        int c[8];
        float g[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = base + 64 * min(k0 + q, w - 1) + lane;
            c[q] = idx[e];
            g[q] = k0 + q < w ? sv[e] : 0.0f;
        }
        double xv[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) xv[q] = x[c[q]];
#pragma unroll
        for (int q = 0; q < 8; ++q) acc += (double)g[q] * xv[q];
    }
    return acc;
}

// The same over packed entries (k_pack_sell): one 32-bit load per entry gives the value and
// the index, row + the signed low half.
__device__ __forceinline__ double sell_row_dot_packed(const uint32_t *__restrict__ pk, int base, int w, int lane,
                                                      int row, const double *__restrict__ x) {
    double acc = 0.0;
    for (int k0 = 0; k0 < w; k0 += 8) {
        uint32_t u[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const uint32_t v = pk[base + 64 * min(k0 + q, w - 1) + lane];
            u[q] = k0 + q < w ? v : (v & 0xffffu);  // past the row's width: value 0, same index
        }
        double xv[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) xv[q] = x[row + (int)(int16_t)(u[q] & 0xffffu)];
#pragma unroll
        for (int q = 0; q < 8; ++q) acc += (double)__uint_as_float(u[q] & 0xffff0000u) * xv[q];
    }
    return acc;
}

// Residual-norm test of the fused flexible PCG's iteration `it` (the reference's stopping rule,
// ConjugateGradientGaussPreSolver.cs:151-153, and the stall fallback), formed identically by
// every workgroup of the preconditioner kernels from the step's partials and the previous
// slot of the state.
struct FcgNorm {
    double rnorm, best;
    int best_it, conv, fallback, restart, trig;
};
__device__ inline FcgNorm fcg_norm_decide(const double *__restrict__ prr, int nstep, int it, int stall, double thr,
                                          const FcgState *__restrict__ pv, double *sm4) {
    FcgNorm d;
    d.rnorm = sqrt(sum_partials(prr, nstep, sm4));
    d.conv = d.rnorm <= thr;
    d.fallback = pv->fallback2;
    d.restart = pv->restart2;
    d.best = pv->best;
    d.best_it = pv->best_it;
    d.trig = 0;
    if (!d.conv && !d.fallback) {
        if (d.rnorm < d.best) {
            d.best = d.rnorm;
            d.best_it = it;
        } else if (it - d.best_it >= stall) {
            d.fallback = 1;
            d.restart = 1;
            d.trig = 1;
        }
    }
    return d;
}

// y = G r for iteration `it` of the fused flexible PCG (skipped once converged or fallen back);
// PK: G's entries packed (BF16 value, 16-bit index difference), else FP32 values and int
// indices
template <bool PK>
__global__ __launch_bounds__(256) void k_fcg_g_sell(const int *__restrict__ sp, const int *__restrict__ scol,
                                                    const float *__restrict__ sv, const uint32_t *__restrict__ pk,
                                                    int n, int nslices, const double *__restrict__ r,
                                                    double *__restrict__ y, const double *__restrict__ prr, int nstep,
                                                    int it, int stall, const FcgState *__restrict__ pv,
                                                    const PcgScalars *__restrict__ S) {
    __shared__ double sm4[4];
    if (S->done) return;
    const FcgNorm d = fcg_norm_decide(prr, nstep, it, stall, S->thr, pv, sm4);
    if (d.conv || d.fallback) return;  // uniform
    const int lane = threadIdx.x & 63, slice = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (slice >= nslices) return;
    const int base = sp[slice], w = (sp[slice + 1] - base) >> 6, row = 64 * slice + lane;
    const double acc = PK ? sell_row_dot_packed(pk, base, w, lane, row, r) : sell_row_dot(scol, sv, base, w, lane, r);
    if (row < n) y[row] = acc;
}

// z = G' y (Jacobi z = D^{-1} r once fallen back) for iteration `it`, with block partials of
// r'z and z'Ap; workgroup 0 records the norm test in slot it & 1 (and the result scalars)
template <bool PK>
__global__ __launch_bounds__(256) void k_fcg_gt_sell(const int *__restrict__ sp, const int *__restrict__ srow,
                                                     const float *__restrict__ sv, const uint32_t *__restrict__ pk,
                                                     int n, int nslices,
                                                     const double *__restrict__ y, double *__restrict__ z,
                                                     const double *__restrict__ r, const double *__restrict__ ap,
                                                     const double *__restrict__ invd, const double *__restrict__ prr,
                                                     int nstep, int it, int stall, const FcgState *__restrict__ pv,
                                                     FcgState *__restrict__ cur, PcgScalars *__restrict__ S,
                                                     double *__restrict__ rz_part, double *__restrict__ zap_part) {
    __shared__ double sm4[4];
    __shared__ double sm[2][4];
    if (S->done) return;
    const FcgNorm d = fcg_norm_decide(prr, nstep, it, stall, S->thr, pv, sm4);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        S->rnorm = d.rnorm;
        S->iters = it;
        if (d.conv) {
            S->done = 1;
        } else {
            cur->rnorm = d.rnorm;
            cur->best = d.best;
            cur->best_it = d.best_it;
            cur->fallback1 = d.fallback;
            cur->restart1 = d.restart;
            S->fallback = d.fallback;
        }
    }
    if (d.conv) return;  // uniform
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double arz = 0.0, azap = 0.0;
    for (int slice = blockIdx.x * 4 + wave; slice < nslices; slice += gridDim.x * 4) {
        const int j = 64 * slice + lane;
        double acc = 0.0;
        if (d.fallback) {
            if (j < n) acc = invd[j] * r[j];
        } else {
            const int base = sp[slice], w = (sp[slice + 1] - base) >> 6;
            acc = PK ? sell_row_dot_packed(pk, base, w, lane, j, y) : sell_row_dot(srow, sv, base, w, lane, y);
        }
        if (j < n) {
            z[j] = acc;
            arz += r[j] * acc;
            azap += acc * ap[j];
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        arz += __shfl_xor(arz, o, 64);
        azap += __shfl_xor(azap, o, 64);
    }
    if (lane == 0) {
        sm[0][wave] = arz;
        sm[1][wave] = azap;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        rz_part[blockIdx.x] = (sm[0][0] + sm[0][1]) + (sm[0][2] + sm[0][3]);
        zap_part[blockIdx.x] = (sm[1][0] + sm[1][1]) + (sm[1][2] + sm[1][3]);
    }
}

// gvt[t] = gv[gtidx[t]]: factor values in the order of the transposed pattern
__global__ void k_gather_t(const int *__restrict__ gtidx, const float *__restrict__ gv, int64_t ng,
                           float *__restrict__ gvt) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ng) gvt[t] = gv[gtidx[t]];
}

// u = G g with g_j = 2 (Az_j - r_j) / (r'e) / ns computed on the fly (scal[1] = r'e)
__global__ __launch_bounds__(256) void k_u_apply(const int *__restrict__ gptr, const int *__restrict__ gcol,
                                                 const float *__restrict__ gv, const double *__restrict__ az,
                                                 const double *__restrict__ r, const double *__restrict__ scal, int ns,
                                                 int n, double *__restrict__ u) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double c = 2.0 / scal[1] / ns;
    double acc = 0.0;
    for (int e = gptr[i]; e < gptr[i + 1]; ++e) {
        const int j = gcol[e];
        acc += (double)gv[e] * (c * (az[j] - r[j]));
    }
    u[i] = acc;
}

// dL/dG_ij += y_i g_j + u_i r_j on the pattern of G
__global__ __launch_bounds__(256) void k_dg_acc(const int *__restrict__ gptr, const int *__restrict__ gcol,
                                                const double *__restrict__ y, const double *__restrict__ u,
                                                const double *__restrict__ az, const double *__restrict__ r,
                                                const double *__restrict__ scal, int ns, int n, float *__restrict__ dG) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double c = 2.0 / scal[1] / ns;
    const double yi = y[i], ui = u[i];
    for (int e = gptr[i]; e < gptr[i + 1]; ++e) {
        const int j = gcol[e];
        dG[e] += (float)(yi * (c * (az[j] - r[j])) + ui * r[j]);
    }
}

// A_kj from the full CSR (rows ascending)
__device__ inline double a_entry(const int *__restrict__ rp, const int *__restrict__ ci, const double *__restrict__ v,
                                 int k, int j) {
    for (int q = rp[k]; q < rp[k + 1]; ++q)
        if (ci[q] == j) return v[q];
    return 0.0;
}

// Kaporin term, stage 1: (G A)_ij on the pattern of G, per-row trace terms and log G_ii,
// fixed-grid partial sums (part[b], part[kRedBlocks + b])
__global__ __launch_bounds__(256) void k_kap_rows(const int *__restrict__ gptr, const int *__restrict__ gcol,
                                                  const float *__restrict__ gv, const int *__restrict__ rp,
                                                  const int *__restrict__ ci, const double *__restrict__ av, int n,
                                                  float *__restrict__ gaE, double *__restrict__ part) {
    __shared__ double sm[2][256];
    double tr = 0.0, ld = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        for (int e = gptr[i]; e < gptr[i + 1]; ++e) {
            const int j = gcol[e];
            double ga = 0.0;
            for (int e2 = gptr[i]; e2 < gptr[i + 1]; ++e2) ga += (double)gv[e2] * a_entry(rp, ci, av, gcol[e2], j);
            gaE[e] = (float)ga;
            tr += (double)gv[e] * ga;
            if (j == i) ld += log((double)gv[e]);
        }
    }
    sm[0][threadIdx.x] = tr;
    sm[1][threadIdx.x] = ld;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            sm[0][threadIdx.x] += sm[0][threadIdx.x + h];
            sm[1][threadIdx.x] += sm[1][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = sm[0][0];
        part[kRedBlocks + blockIdx.x] = sm[1][0];
    }
}

// stage 2: scal[4] = tr(G A G'), scal[2] += kappa (log(T / n) - (2/n) sum log G_ii)
__global__ __launch_bounds__(256) void k_kap_final(const double *__restrict__ part, int n, double kappa,
                                                   double *__restrict__ scal) {
    __shared__ double sm[2][256];
    double a = 0.0, b = 0.0;
    for (int q = threadIdx.x; q < kRedBlocks; q += 256) {
        a += part[q];
        b += part[kRedBlocks + q];
    }
    sm[0][threadIdx.x] = a;
    sm[1][threadIdx.x] = b;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            sm[0][threadIdx.x] += sm[0][threadIdx.x + h];
            sm[1][threadIdx.x] += sm[1][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        scal[4] = sm[0][0];
        scal[2] += kappa * (log(sm[0][0] / n) - 2.0 / n * sm[1][0]);
    }
}

// stage 3: dL_K/dG_ij = kappa ((2/T) (G A)_ij - (2/n) [j == i] / G_ii)
__global__ __launch_bounds__(256) void k_kap_grad(const int *__restrict__ gptr, const int *__restrict__ gcol,
                                                  const float *__restrict__ gv, const float *__restrict__ gaE,
                                                  const double *__restrict__ scal, double kappa, int n,
                                                  float *__restrict__ dG) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double T = scal[4];
    for (int e = gptr[i]; e < gptr[i + 1]; ++e)
        dG[e] += (float)(kappa * (2.0 / T * (double)gaE[e] - (gcol[e] == i ? 2.0 / n / (double)gv[e] : 0.0)));
}

// Backward of the readouts: per node (32 lanes), the entries of its row (node as i) and of its
// column (node as j, transposed pattern); weight gradients per block into part.
__global__ __launch_bounds__(256) void k_mp_readout_bwd(const float *__restrict__ h3, const float *__restrict__ w,
                                                        const int *__restrict__ gptr, const int *__restrict__ gcol,
                                                        const int *__restrict__ gtptr, const int *__restrict__ gtrow,
                                                        const int *__restrict__ gtidx, const float *__restrict__ ge1,
                                                        const float *__restrict__ ge2, const float *__restrict__ gv,
                                                        const float *__restrict__ dG, const double *__restrict__ diag,
                                                        int n, float *__restrict__ dH, float *__restrict__ part) {
    __shared__ float sm[kNB][W];
    const int g = threadIdx.x / W, f = threadIdx.x % W;
    const float *rd = w + kMpOffT;  // w_t, u1, v1, w1, u2, v2, w2
    float acc[7];
#pragma unroll
    for (int q = 0; q < 7; ++q) acc[q] = 0.0f;
    for (int base = blockIdx.x * kNB; base < n; base += gridDim.x * kNB) {
        const int i = base + g;
        if (i >= n) continue;
        const float hi = h3[(int64_t)i * W + f];
        float dh = 0.0f;
        for (int e = gptr[i]; e < gptr[i + 1]; ++e) {
            const int j = gcol[e];
            if (j == i) {
                const float dt = dG[e] * gv[e];
                acc[0] += dt * hi;
                dh += dt * rd[f];
                continue;
            }
            const float hj = h3[(int64_t)j * W + f];
            const float sj = (float)(1.0 / sqrt(diag[j]));
            const float d1 = dG[e] * ge1[e] * sj, d2 = dG[e] * ge2[e] * sj;
            acc[1] += d1 * hi * hj;
            acc[2] += d1 * hi;
            acc[3] += d1 * hj;
            acc[4] += d2 * hi * hj;
            acc[5] += d2 * hi;
            acc[6] += d2 * hj;
            dh += d1 * (rd[W + f] * hj + rd[2 * W + f]) + d2 * (rd[4 * W + f] * hj + rd[5 * W + f]);
        }
        const float si = (float)(1.0 / sqrt(diag[i]));
        for (int t = gtptr[i]; t < gtptr[i + 1]; ++t) {
            const int k = gtrow[t];  // entry (k, i) with k > i
            if (k == i) continue;
            const int e = gtidx[t];
            const float hk = h3[(int64_t)k * W + f];
            const float d1 = dG[e] * ge1[e] * si, d2 = dG[e] * ge2[e] * si;
            dh += d1 * (rd[W + f] * hk + rd[3 * W + f]) + d2 * (rd[4 * W + f] * hk + rd[6 * W + f]);
        }
        dH[(int64_t)i * W + f] = dh;
    }
    float *dst = part + (int64_t)blockIdx.x * kMpP + kMpOffT;
    for (int q = 0; q < 7; ++q) group_reduce_add(acc[q], sm, dst + q * W);
}

__global__ __launch_bounds__(256) void k_mp_in_bwd(const float *__restrict__ x, const float *__restrict__ dH0, int n,
                                                   float *__restrict__ part) {
    __shared__ float sm[kNB][W];
    const int g = threadIdx.x / W, f = threadIdx.x % W;
    float a[3] = {0.0f, 0.0f, 0.0f};
    for (int base = blockIdx.x * kNB; base < n; base += gridDim.x * kNB) {
        const int i = base + g;
        if (i >= n) continue;
        const float d = dH0[(int64_t)i * W + f];
        for (int c = 0; c < 3; ++c) a[c] += d * x[3 * i + c];
    }
    // W_in[f][c] at 3f + c: reduce each c, then scatter
    for (int c = 0; c < 3; ++c) {
        sm[g][f] = a[c];
        __syncthreads();
        if (g == 0) {
            float t = 0.0f;
            for (int q = 0; q < kNB; ++q) t += sm[q][f];
            part[(int64_t)blockIdx.x * kMpP + 3 * f + c] += t;
        }
        __syncthreads();
    }
}

// BF16 copy: for layer l, output block nb (features 16nb..16nb+15), k chunk kk, lane, i:
// B[k][col] = [W1_l | W2_l][col][k] with k = 16kk + 4(lane/16) + i, col = 16nb + lane%16
// (the B-operand layout of v_mfma_f32_16x16x16_bf16); then W_in and the readouts.
__global__ void k_mp_pack(const float *__restrict__ w, uint16_t *__restrict__ wb) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < kMpBf16Vec) {
        const int i = idx & 3, lane = (idx >> 2) & 63, kk = (idx >> 8) & 3, nb = (idx >> 10) & 1, l = idx >> 11;
        const int k = 16 * kk + 4 * (lane >> 4) + i, col = 16 * nb + (lane & 15);
        const float *wl = w + kMpOffW + l * 2 * W * W;
        const float v = k < W ? wl[col * W + k] : wl[W * W + col * W + (k - W)];
        wb[idx] = f2bf(v);
    } else if (idx < kMpBf16Vec + 3 * W + 7 * W) {
        const int t = idx - kMpBf16Vec;  // W_in (96), then the readouts (224)
        wb[idx] = f2bf(t < 3 * W ? w[t] : w[kMpOffT + (t - 3 * W)]);
    }
}

__global__ __launch_bounds__(256) void k_mpb_input(const float *__restrict__ x, const uint16_t *__restrict__ wv, int n,
                                                   uint16_t *__restrict__ H0) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int i = (int)(e / W), f = (int)(e % W);
    if (i >= n) return;
    H0[e] = f2bf(bf2f(wv[3 * f]) * x[3 * i] + bf2f(wv[3 * f + 1]) * x[3 * i + 1] + bf2f(wv[3 * f + 2]) * x[3 * i + 2]);
}

void ensure_ws(Mpnn &net, int n, int64_t ng) {
    if (!net.part.n) {
        net.part.alloc((size_t)kMpBlocks * kMpP);
        net.dpart.alloc(2 * kRedBlocks + 3 * kRedBlocks);
        net.scal.alloc(8);
    }
    if (net.ws_n >= n && net.ws_ng >= ng) return;
    const size_t nw = (size_t)std::max(n, 1) * W, nn = std::max(n, 1), gg = (size_t)std::max<int64_t>(ng, 1);
    net.H.alloc(4 * nw);
    net.Mb.alloc(3 * nw);
    net.Pre.alloc(3 * nw);
    net.dH.alloc(nw);
    net.dHp.alloc(nw);
    net.dM.alloc(nw);
    net.hb0.alloc(nw);
    net.hb1.alloc(nw);
    net.gvf.alloc(gg);
    net.dG.alloc(gg);
    net.gaE.alloc(gg);
    net.gv.alloc(gg);
    net.gvt.alloc(gg);
    net.gv_graph = nullptr;
    net.y.alloc(nn);
    net.z.alloc(nn);
    net.az.alloc(nn);
    net.u.alloc(nn);
    net.ws_n = n;
    net.ws_ng = ng;
}

}  // namespace

// ---- host side -------------------------------------------------------------------

// Graph data of the learned factor, built on the host from the full CSR in the same order
// as the oracle's restatement (rows ascending; the factor pattern of row i is i, its lower
// neighbours and their neighbours below i, sorted).
void mp_graph_prepare(Matrix &A, hipStream_t st) {
    if (A.mp && A.mp->n == A.n) return;
    if (A.diag_status)
        throw Error(MG_E_NOT_PD, "The diagonal element is zero or invalid while building the learned preconditioner.");
    const int n = A.n;
    const int64_t nnz = A.nnz_full;
    std::vector<int> rp(n + 1), ci(std::max<int64_t>(nnz, 1));
    std::vector<double> av(std::max<int64_t>(nnz, 1));
    MG_HIP(hipStreamSynchronize(st));
    MG_HIP(hipMemcpy(rp.data(), A.rowptr.p, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
    if (nnz) {
        MG_HIP(hipMemcpy(ci.data(), A.colidx.p, sizeof(int) * nnz, hipMemcpyDeviceToHost));
        MG_HIP(hipMemcpy(av.data(), A.vals.p, sizeof(double) * nnz, hipMemcpyDeviceToHost));
    }
    std::vector<double> d(n, 0.0);
    double mean = 0.0;
    for (int i = 0; i < n; ++i) {
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] == i) d[i] = av[k];
        mean += d[i];
    }
    mean = n ? mean / n : 1.0;
    std::vector<float> ew(std::max<int64_t>(nnz, 1), 0.0f), x(3 * (size_t)std::max(n, 1));
    std::vector<double> e1(std::max<int64_t>(nnz, 1), 0.0);
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            const int j = ci[k];
            if (j == i) continue;
            e1[k] = av[k] / std::sqrt(d[i] * d[j]);
            ew[k] = (float)e1[k];
            s += std::fabs(e1[k]);
        }
        x[3 * i] = 1.0f;
        x[3 * i + 1] = (float)std::log(d[i] / mean);
        x[3 * i + 2] = (float)s;
    }
    std::vector<int> gptr(n + 1, 0), gcol;
    std::vector<float> ge1, ge2;
    std::vector<double> w1(std::max(n, 1), 0.0), w2(std::max(n, 1), 0.0);
    std::vector<char> seen(std::max(n, 1), 0);
    std::vector<int> cols;
    for (int i = 0; i < n; ++i) {
        cols.clear();
        seen[i] = 1;
        cols.push_back(i);
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            const int j = ci[k];
            if (j == i) continue;
            if (j < i) {
                w1[j] += e1[k];
                if (!seen[j]) { seen[j] = 1; cols.push_back(j); }
            }
            for (int q = rp[j]; q < rp[j + 1]; ++q) {
                const int l = ci[q];
                if (l == j || l >= i) continue;
                w2[l] += e1[k] * e1[q];
                if (!seen[l]) { seen[l] = 1; cols.push_back(l); }
            }
        }
        std::sort(cols.begin(), cols.end());
        for (int j : cols) {
            gcol.push_back(j);
            ge1.push_back((float)w1[j]);
            ge2.push_back((float)w2[j]);
            w1[j] = w2[j] = 0.0;
            seen[j] = 0;
        }
        gptr[i + 1] = (int)gcol.size();
    }
    const int64_t ng = (int64_t)gcol.size();
    std::vector<int> gtptr(n + 1, 0), gtrow(std::max<int64_t>(ng, 1)), gtidx(std::max<int64_t>(ng, 1));
    for (int64_t e = 0; e < ng; ++e) ++gtptr[gcol[e] + 1];
    for (int j = 0; j < n; ++j) gtptr[j + 1] += gtptr[j];
    {
        std::vector<int> at(gtptr.begin(), gtptr.end() - 1);
        for (int i = 0; i < n; ++i)
            for (int e = gptr[i]; e < gptr[i + 1]; ++e) {
                const int t = at[gcol[e]]++;
                gtrow[t] = i;
                gtidx[t] = e;
            }
    }
    auto g = std::make_shared<MpGraph>();
    {
        // sliced ELL of G (entry e's value at gv[e]) and of G' (value at gv[gtidx[t]])
        const int nsl = (n + 63) / 64;
        bool fits = true;  // every index difference fits 16 bits
        auto sell = [&](const std::vector<int> &ptr, const std::vector<int> &idx, const int *vmap,
                        DevArray<int> &dptr, DevArray<int> &didx, DevArray<int> &dsrc, DevArray<uint16_t> &ddl) {
            std::vector<int> sp(nsl + 1, 0);
            for (int sl = 0; sl < nsl; ++sl) {
                int w = 0;
                for (int r = 64 * sl; r < std::min(n, 64 * sl + 64); ++r) w = std::max(w, ptr[r + 1] - ptr[r]);
                sp[sl + 1] = sp[sl] + 64 * w;
            }
            if ((double)sp[nsl] > 1.5 * (double)std::max<int64_t>(ng, 1)) return false;
            std::vector<int> si(std::max(sp[nsl], 1)), ss(std::max(sp[nsl], 1), -1);
            std::vector<uint16_t> dl(std::max(sp[nsl], 1), 0);
            for (int sl = 0; sl < nsl; ++sl) {
                const int w = (sp[sl + 1] - sp[sl]) / 64;
                for (int l = 0; l < 64; ++l) {
                    const int r = std::min(64 * sl + l, n - 1);
                    const int len = 64 * sl + l < n ? ptr[r + 1] - ptr[r] : 0;
                    for (int k = 0; k < w; ++k) {
                        const int e = sp[sl] + 64 * k + l;
                        si[e] = k < len ? idx[ptr[r] + k] : r;
                        ss[e] = k < len ? (vmap ? vmap[ptr[r] + k] : ptr[r] + k) : -1;
                        const int d = si[e] - (64 * sl + l);  // from the slot's row (past n too)
                        fits = fits && d >= -32768 && d <= 32767;
                        dl[e] = (uint16_t)(int16_t)(fits ? d : 0);
                    }
                }
            }
            dptr.upload(sp);
            didx.upload(si);
            dsrc.upload(ss);
            ddl.upload(dl);
            return true;
        };
        if (n > 0 && sell(gptr, gcol, nullptr, g->gs_ptr, g->gs_col, g->gs_src, g->gs_dl) &&
            sell(gtptr, gtrow, gtidx.data(), g->gts_ptr, g->gts_row, g->gts_src, g->gts_dl)) {
            g->nslices = nsl;
            g->packed = fits;
        }
    }
    g->n = n;
    g->ng = ng;
    g->ew.upload(ew);
    g->x.upload(x);
    g->diag.upload(d);
    g->gptr.upload(gptr);
    g->gcol.upload(gcol);
    g->ge1.upload(ge1);
    g->ge2.upload(ge2);
    g->gtptr.upload(gtptr);
    g->gtrow.upload(gtrow);
    g->gtidx.upload(gtidx);
    A.mp = g;
}

// Deterministic initialisation: W_in ~ N(0, 1), W1, W2 ~ N(0, 1/(2*32)), readouts ~ N(0, 0.01^2)
// (small readouts start the factor near Jacobi).
void mpnn_init(Mpnn &net, uint64_t seed, hipStream_t st) {
    std::mt19937_64 gen(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::vector<float> w(kMpP);
    for (int p = 0; p < kMpP; ++p) {
        double sd = 1.0;
        if (p >= kMpOffW && p < kMpOffT) sd = 1.0 / std::sqrt(2.0 * W);
        else if (p >= kMpOffT) sd = 0.01;
        w[p] = (float)(nd(gen) * sd);
    }
    net.w.upload(w);
    std::vector<float> zero(kMpP, 0.0f);
    net.grad.upload(zero);
    net.am.upload(zero);
    net.av.upload(zero);
    net.adam_t = 0;
    net.wb.alloc(kMpBf16Vec + 10 * W);
    mpnn_pack_bf16(net, st);
}

void mpnn_pack_bf16(Mpnn &net, hipStream_t st) {
    const int total = kMpBf16Vec + 10 * W;
    hipLaunchKernelGGL(k_mp_pack, dim3(nblk(total, 256)), dim3(256), 0, st, net.w.p, net.wb.p);
    MG_HIP(hipGetLastError());
    net.gv_graph = nullptr;  // factor values are stale
}

namespace {

// FP32 forward of the network and the factor values (gvf), keeping the activations
void forward_f32(Mpnn &net, Matrix &A, hipStream_t st) {
    const int n = A.n;
    const MpGraph &g = *A.mp;
    const int64_t nw = (int64_t)n * W;
    hipLaunchKernelGGL(k_mp_input, dim3(nblk(nw, 256)), dim3(256), 0, st, g.x.p, net.w.p, n, net.H.p);
    for (int l = 0; l < 3; ++l)
        hipLaunchKernelGGL(k_mp_layer, dim3(kMpBlocks), dim3(256), 0, st, net.w.p + kMpOffW + l * 2 * W * W,
                           A.rowptr.p, A.colidx.p, g.ew.p, n, net.H.p + l * nw, net.Mb.p + l * nw, net.Pre.p + l * nw,
                           net.H.p + (l + 1) * nw, l < 2 ? 1 : 0);
    hipLaunchKernelGGL((k_mp_readout<float, float, float>), dim3(kMpBlocks), dim3(256), 0, st, net.H.p + 3 * nw,
                       net.w.p + kMpOffT, g.gptr.p, g.gcol.p, g.ge1.p, g.ge2.p, g.diag.p, n, net.gvf.p);
}

void begin_batch(Mpnn &net, Matrix &A, hipStream_t st) {
    const int64_t cnt = (int64_t)kMpBlocks * kMpP;
    hipLaunchKernelGGL(k_fill_f32, dim3(nblk(cnt, 256)), dim3(256), 0, st, net.part.p, cnt, 0.0f);
    const int64_t ng = A.mp->ng;
    if (ng) hipLaunchKernelGGL(k_fill_f32, dim3(nblk(ng, 256)), dim3(256), 0, st, net.dG.p, ng, 0.0f);
    hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(3), 0, st, net.scal.p, 3, 0.0);
}

// pair term of one sample: loss into scal[2], dL/dG into dG
void sample_grad(Mpnn &net, Matrix &A, const double *d_r, const double *d_e, int ns, hipStream_t st) {
    const int n = A.n;
    const MpGraph &g = *A.mp;
    const dim3 gn(nblk(n, 256)), gr8(nblk((int64_t)n * kRowLanes, 256));
    hipLaunchKernelGGL((k_g_apply<float>), gr8, dim3(256), 0, st, g.gptr.p, g.gcol.p, net.gvf.p, d_r, n, net.y.p,
                       (const PcgScalars *)nullptr);
    hipLaunchKernelGGL((k_gt_apply<float>), gr8, dim3(256), 0, st, g.gtptr.p, g.gtrow.p, g.gtidx.p,
                       net.gvf.p, (const float *)nullptr, net.y.p, n, net.z.p, (const double *)nullptr,
                       (const double *)nullptr, (double *)nullptr, (double *)nullptr, (const double *)nullptr,
                       (const PcgScalars *)nullptr);
    spmv_device(A, 1, net.z.p, n, net.az.p, n, st);
    hipLaunchKernelGGL(k_mp_loss_part, dim3(kRedBlocks), dim3(256), 0, st, net.z.p, net.az.p, d_r, d_e, n,
                       net.dpart.p);
    hipLaunchKernelGGL(k_mp_loss_final, dim3(1), dim3(256), 0, st, net.dpart.p, ns, net.scal.p);
    hipLaunchKernelGGL(k_u_apply, gn, dim3(256), 0, st, g.gptr.p, g.gcol.p, net.gvf.p, net.az.p, d_r, net.scal.p, ns,
                       n, net.u.p);
    hipLaunchKernelGGL(k_dg_acc, gn, dim3(256), 0, st, g.gptr.p, g.gcol.p, net.y.p, net.u.p, net.az.p, d_r,
                       net.scal.p, ns, n, net.dG.p);
}

// Kaporin term, readout and network backward, gradient reduction (after the samples of a batch)
void finish_batch(Mpnn &net, Matrix &A, double kappa, hipStream_t st) {
    const int n = A.n;
    const MpGraph &g = *A.mp;
    const int64_t nw = (int64_t)n * W;
    if (kappa != 0.0 && n > 0) {
        hipLaunchKernelGGL(k_kap_rows, dim3(kRedBlocks), dim3(256), 0, st, g.gptr.p, g.gcol.p, net.gvf.p, A.rowptr.p,
                           A.colidx.p, A.vals.p, n, net.gaE.p, net.dpart.p);
        hipLaunchKernelGGL(k_kap_final, dim3(1), dim3(256), 0, st, net.dpart.p, n, kappa, net.scal.p);
        hipLaunchKernelGGL(k_kap_grad, dim3(nblk(n, 256)), dim3(256), 0, st, g.gptr.p, g.gcol.p, net.gvf.p, net.gaE.p,
                           net.scal.p, kappa, n, net.dG.p);
    }
    hipLaunchKernelGGL(k_mp_readout_bwd, dim3(kMpBlocks), dim3(256), 0, st, net.H.p + 3 * nw, net.w.p, g.gptr.p,
                       g.gcol.p, g.gtptr.p, g.gtrow.p, g.gtidx.p, g.ge1.p, g.ge2.p, net.gvf.p, net.dG.p, g.diag.p, n,
                       net.dH.p, net.part.p);
    float *dh = net.dH.p, *dhp = net.dHp.p;
    for (int l = 2; l >= 0; --l) {
        hipLaunchKernelGGL(k_mp_layer_bwd, dim3(kMpBlocks), dim3(256), 0, st, net.w.p + kMpOffW + l * 2 * W * W, n,
                           dh, net.Pre.p + l * nw, net.H.p + l * nw, net.Mb.p + l * nw, l < 2 ? 1 : 0, dhp, net.dM.p,
                           net.part.p + kMpOffW + l * 2 * W * W);
        hipLaunchKernelGGL(k_mp_agg_bwd, dim3(nblk(nw, 256)), dim3(256), 0, st, A.rowptr.p, A.colidx.p, g.ew.p, n,
                           net.dM.p, dhp);
        std::swap(dh, dhp);
    }
    hipLaunchKernelGGL(k_mp_in_bwd, dim3(kMpBlocks), dim3(256), 0, st, g.x.p, dh, n, net.part.p);
    hipLaunchKernelGGL(k_mp_grad_reduce, dim3(nblk(kMpP, 256)), dim3(256), 0, st, net.part.p, kMpBlocks, net.grad.p);
}

__global__ void k_mp_epoch_acc(double *__restrict__ scal, double weight) { scal[3] += scal[2] * weight; }

}  // namespace

void mpnn_apply(Mpnn &net, Matrix &A, const double *d_r, double *d_z, int precision, hipStream_t st) {
    mp_graph_prepare(A, st);
    ensure_ws(net, A.n, A.mp->ng);
    const int n = A.n;
    if (n == 0) return;
    const MpGraph &g = *A.mp;
    if (precision == MG_BF16) {
        mpnn_factor_bf16(net, A, st);
        hipLaunchKernelGGL((k_g_apply<float>), dim3(nblk((int64_t)n * kRowLanes, 256)), dim3(256), 0, st, g.gptr.p,
                           g.gcol.p, net.gv.p, d_r, n, net.y.p, (const PcgScalars *)nullptr);
        hipLaunchKernelGGL((k_gt_apply<float>), dim3(nblk((int64_t)n * kRowLanes, 256)), dim3(256), 0, st, g.gtptr.p, g.gtrow.p,
                           g.gtidx.p, net.gv.p, net.gvt.p, net.y.p, n, d_z, (const double *)nullptr,
                           (const double *)nullptr, (double *)nullptr, (double *)nullptr, (const double *)nullptr,
                           (const PcgScalars *)nullptr);
    } else {
        forward_f32(net, A, st);
        hipLaunchKernelGGL((k_g_apply<float>), dim3(nblk((int64_t)n * kRowLanes, 256)), dim3(256), 0, st, g.gptr.p,
                           g.gcol.p, net.gvf.p, d_r, n, net.y.p, (const PcgScalars *)nullptr);
        hipLaunchKernelGGL((k_gt_apply<float>), dim3(nblk((int64_t)n * kRowLanes, 256)), dim3(256), 0, st, g.gtptr.p, g.gtrow.p,
                           g.gtidx.p, net.gvf.p, (const float *)nullptr, net.y.p, n, d_z, (const double *)nullptr,
                           (const double *)nullptr, (double *)nullptr, (double *)nullptr, (const double *)nullptr,
                           (const PcgScalars *)nullptr);
    }
    MG_HIP(hipGetLastError());
}

void mpnn_factor_bf16(Mpnn &net, Matrix &A, hipStream_t st) {
    mp_graph_prepare(A, st);
    ensure_ws(net, A.n, A.mp->ng);
    const MpGraph &g = *A.mp;
    if (net.gv_graph == &g) return;
    const int n = A.n;
    if (n == 0) return;
    const int64_t nw = (int64_t)n * W;
    const uint16_t *wv = net.wb.p + kMpBf16Vec;
    hipLaunchKernelGGL(k_mpb_input, dim3(nblk(nw, 256)), dim3(256), 0, st, g.x.p, wv, n, net.hb0.p);
    uint16_t *hin = net.hb0.p, *hout = net.hb1.p;
    const int tiles = (n + 15) / 16;
    const int grid = std::min(nblk(tiles, 4), 2048);
    for (int l = 0; l < 3; ++l) {
        hipLaunchKernelGGL(k_mpb_layer, dim3(grid), dim3(256), 0, st, net.wb.p + l * kMpBf16Layer, A.rowptr.p,
                           A.colidx.p, g.ew.p, n, hin, hout, l < 2 ? 1 : 0, (const PcgScalars *)nullptr);
        std::swap(hin, hout);
    }
    hipLaunchKernelGGL((k_mp_readout<uint16_t, uint16_t, float>), dim3(kMpBlocks), dim3(256), 0, st, hin,
                       wv + 3 * W, g.gptr.p, g.gcol.p, g.ge1.p, g.ge2.p, g.diag.p, n, net.gv.p);
    if (g.ng)
        hipLaunchKernelGGL(k_gather_t, dim3(nblk(g.ng, 256)), dim3(256), 0, st, g.gtidx.p, net.gv.p, g.ng, net.gvt.p);
    if (g.nslices) {
        if (net.gvs.n != g.gs_src.n) net.gvs.alloc(g.gs_src.n);
        if (net.gvts.n != g.gts_src.n) net.gvts.alloc(g.gts_src.n);
        hipLaunchKernelGGL(k_gather_sell, dim3(nblk((int64_t)g.gs_src.n, 256)), dim3(256), 0, st, g.gs_src.p,
                           (int64_t)g.gs_src.n, net.gv.p, net.gvs.p);
        hipLaunchKernelGGL(k_gather_sell, dim3(nblk((int64_t)g.gts_src.n, 256)), dim3(256), 0, st, g.gts_src.p,
                           (int64_t)g.gts_src.n, net.gv.p, net.gvts.p);
        if (g.packed) {
            if (net.gps.n != g.gs_src.n) net.gps.alloc(g.gs_src.n);
            if (net.gpts.n != g.gts_src.n) net.gpts.alloc(g.gts_src.n);
            hipLaunchKernelGGL(k_pack_sell, dim3(nblk((int64_t)g.gs_src.n, 256)), dim3(256), 0, st, g.gs_src.p,
                               g.gs_dl.p, (int64_t)g.gs_src.n, net.gv.p, net.gps.p);
            hipLaunchKernelGGL(k_pack_sell, dim3(nblk((int64_t)g.gts_src.n, 256)), dim3(256), 0, st, g.gts_src.p,
                               g.gts_dl.p, (int64_t)g.gts_src.n, net.gv.p, net.gpts.p);
        }
    }
    MG_HIP(hipGetLastError());
    net.gv_graph = &g;
}

void mpnn_apply_factor(Mpnn &net, Matrix &A, const double *d_r, double *d_z, const PcgScalars *S, const double *d_ap,
                       double *rz_part, double *zap_part, hipStream_t st) {
    const MpGraph &g = *A.mp;
    const int n = A.n;
    if (g.nslices && net.gv_graph == &g && net.gvs.n == g.gs_src.n) {
        hipLaunchKernelGGL(k_g_apply_sell, dim3((g.nslices + 3) / 4), dim3(256), 0, st, g.gs_ptr.p, g.gs_col.p,
                           net.gvs.p, n, g.nslices, d_r, net.y.p, S);
        hipLaunchKernelGGL(k_gt_apply_sell, dim3(mpnn_partial_blocks(n)), dim3(256), 0, st, g.gts_ptr.p, g.gts_row.p,
                           net.gvts.p, n, g.nslices, net.y.p, d_z, d_r, d_ap, rz_part, zap_part, A.invdiag.p, S);
        return;
    }
    // the transposed product runs on a fixed grid (one partial sum per block; after a fallback it
    // computes the Jacobi z); the row product is skipped after a fallback
    hipLaunchKernelGGL((k_g_apply<float>), dim3(nblk((int64_t)n * kRowLanes, 256)), dim3(256), 0, st, g.gptr.p,
                       g.gcol.p, net.gv.p, d_r, n, net.y.p, S);
    hipLaunchKernelGGL((k_gt_apply<float>), dim3(mpnn_partial_blocks(n)), dim3(256), 0, st, g.gtptr.p, g.gtrow.p,
                       g.gtidx.p, net.gv.p, net.gvt.p, net.y.p, n, d_z, d_r, d_ap, rz_part, zap_part, A.invdiag.p, S);
}

int mpnn_fcg_partials(const Matrix &A) { return A.mp ? std::max(1, (A.mp->nslices + 3) / 4) : 1; }

bool mpnn_fcg_apply(Mpnn &net, Matrix &A, const double *d_r, double *d_z, const double *d_ap, const double *prr,
                    int nstep, int it, int stall, const FcgState *fs_prev, FcgState *fs_cur, PcgScalars *S,
                    double *rz_part, double *zap_part, hipStream_t st) {
    const MpGraph &g = *A.mp;
    if (!g.nslices || net.gv_graph != &g || net.gvs.n != g.gs_src.n) return false;
    const int grid = mpnn_fcg_partials(A);
    const bool pk = g.packed && net.gps.n == g.gs_src.n && net.gpts.n == g.gts_src.n;
    hipLaunchKernelGGL(pk ? k_fcg_g_sell<true> : k_fcg_g_sell<false>, dim3(grid), dim3(256), 0, st, g.gs_ptr.p,
                       g.gs_col.p, net.gvs.p, net.gps.p, A.n, g.nslices, d_r, net.y.p, prr, nstep, it, stall, fs_prev,
                       S);
    hipLaunchKernelGGL(pk ? k_fcg_gt_sell<true> : k_fcg_gt_sell<false>, dim3(grid), dim3(256), 0, st, g.gts_ptr.p,
                       g.gts_row.p, net.gvts.p, net.gpts.p, A.n, g.nslices, net.y.p, d_z, d_r, d_ap, A.invdiag.p, prr,
                       nstep, it, stall, fs_prev, fs_cur, S, rz_part, zap_part);
    return true;
}

double mpnn_loss_grad(Mpnn &net, Matrix &A, int ns, const double *d_R, const double *d_E, int64_t ld, double kappa,
                      hipStream_t st) {
    mp_graph_prepare(A, st);
    ensure_ws(net, A.n, A.mp->ng);
    if (A.n == 0) return 0.0;
    begin_batch(net, A, st);
    forward_f32(net, A, st);
    for (int t = 0; t < ns; ++t) sample_grad(net, A, d_R + t * ld, d_E + t * ld, ns, st);
    finish_batch(net, A, kappa, st);
    MG_HIP(hipGetLastError());
    double scal[8];
    MG_HIP(hipMemcpyAsync(scal, net.scal.p, sizeof(scal), hipMemcpyDeviceToHost, st));
    MG_HIP(hipStreamSynchronize(st));
    return scal[2];
}

// Mini-batch Adam over the ns pairs (columns of R and E), `epochs` passes in a seeded
// shuffled order; losses[e] = mean training loss of epoch e (before each batch's update).
void mpnn_train(Mpnn &net, Matrix &A, int ns, const double *d_R, const double *d_E, int64_t ld, int epochs, int batch,
                float lr, double kappa, uint64_t seed, double *losses, hipStream_t st) {
    if (ns <= 0 || epochs <= 0) return;
    if (batch <= 0) batch = 16;
    mp_graph_prepare(A, st);
    ensure_ws(net, A.n, A.mp->ng);
    if (A.n == 0) return;
    std::mt19937_64 gen(seed);
    std::vector<int> order(ns);
    for (int t = 0; t < ns; ++t) order[t] = t;
    for (int e = 0; e < epochs; ++e) {
        std::shuffle(order.begin(), order.end(), gen);
        hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(8), 0, st, net.scal.p, 8, 0.0);
        for (int b0 = 0; b0 < ns; b0 += batch) {
            const int nb = std::min(batch, ns - b0);
            begin_batch(net, A, st);
            forward_f32(net, A, st);
            for (int t = 0; t < nb; ++t) {
                const int s = order[b0 + t];
                sample_grad(net, A, d_R + s * ld, d_E + s * ld, nb, st);
            }
            finish_batch(net, A, kappa, st);
            hipLaunchKernelGGL(k_mp_epoch_acc, dim3(1), dim3(1), 0, st, net.scal.p, (double)nb / ns);
            mpnn_adam_step(net, lr, st);
        }
        double scal[8];
        MG_HIP(hipMemcpyAsync(scal, net.scal.p, sizeof(scal), hipMemcpyDeviceToHost, st));
        MG_HIP(hipStreamSynchronize(st));
        if (losses) losses[e] = scal[3];
    }
    mpnn_pack_bf16(net, st);
    MG_HIP(hipGetLastError());
}

void mpnn_adam_step(Mpnn &net, float lr, hipStream_t st) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    net.adam_t += 1;
    const float c1 = 1.0f - std::pow(b1, (float)net.adam_t), c2 = 1.0f - std::pow(b2, (float)net.adam_t);
    hipLaunchKernelGGL(k_mp_adam, dim3(nblk(kMpP, 256)), dim3(256), 0, st, net.w.p, net.grad.p, net.am.p, net.av.p,
                       lr, b1, b2, eps, c1, c2);
    MG_HIP(hipGetLastError());
}

namespace {
// one MFMA with the layer kernel's operand layout: lane l supplies a[l%16][4(l/16)+i] and
// b[4(l/16)+i][l%16]; it receives d[4(l/16)+i][l%16]
__global__ void k_mfma_bf16_probe(const float *__restrict__ a, const float *__restrict__ b, float *__restrict__ d) {
    const int lane = threadIdx.x;
    mp_s4 av, bv;
    for (int i = 0; i < 4; ++i) {
        const int k = 4 * (lane >> 4) + i;
        av[i] = (short)f2bf(a[(lane & 15) * 16 + k]);
        bv[i] = (short)f2bf(b[k * 16 + (lane & 15)]);
    }
    mp_f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    for (int i = 0; i < 4; ++i) d[(4 * (lane >> 4) + i) * 16 + (lane & 15)] = acc[i];
}
}  // namespace

void selftest_mfma_bf16(const float *a, const float *b, float *d) {
    DevArray<float> da(256), db(256), dd(256);
    MG_HIP(hipMemcpy(da.p, a, sizeof(float) * 256, hipMemcpyHostToDevice));
    MG_HIP(hipMemcpy(db.p, b, sizeof(float) * 256, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mfma_bf16_probe, dim3(1), dim3(64), 0, 0, da.p, db.p, dd.p);
    MG_HIP(hipGetLastError());
    MG_HIP(hipMemcpy(d, dd.p, sizeof(float) * 256, hipMemcpyDeviceToHost));
}

int mpnn_partial_blocks(int n) { return std::max(1, std::min(nblk((int64_t)n * kRowLanes, 256), 2048)); }

}  // namespace mg
