// mpnn.hip -- learned preconditioner: a small message-passing network on the matrix
// graph, used as a correction on top of CGGP's Jacobi preconditioner
// (CGGP/JacobiPreconditioner.cs:12-56) inside a flexible PCG.  MAGLA has no
// counterpart; the network is defined here:
//
//   inputs   u_i = r_i / s, s = ||r|| / sqrt(n);  c_i = log(a_ii / mean a_ii)
//   edges    ew_ij = a_ij / sqrt(a_ii a_jj)  (j != i, full symmetric pattern)
//   h0_i   = u_i (w_e + c_i w_c)                                  (width 32)
//   layer l: m_i = sum_j ew_ij h_j ;  h'_i = act(W1_l h_i + W2_l m_i)   (l = 0, 1, 2;
//            act = ReLU for l < 2, identity for l = 2)
//   output  y_i = w_d . h3_i ;  z_i = (r_i + s y_i) / a_ii
//
// No biases and ReLU keep z(alpha r) = alpha z(r) for alpha > 0.  Training
// minimises the A-norm error of z against the exact error e = A^{-1} r of
// each (r, e) pair, L = (z - e)'A(z - e) / (e'Ae) = (z'Az - 2 z'r) / (r'e) + 1,
// so dL/dz = 2 (Az - r) / (r'e).  Training runs in FP32 (SIMT kernels, Adam);
// inference in BF16 weights and activations with FP32 accumulation on MFMA
// (v_mfma_f32_16x16x16_bf16: per 16-node tile, [h | m] (16 x 64) x [W1 | W2]^T
// (64 x 32)).  Every reduction has a fixed grid and order.
#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

#include "device.h"

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

// ---- graph data ------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_mp_diag(const int *__restrict__ dpos, const double *__restrict__ vals, int n,
                                                 double *__restrict__ diag, double *__restrict__ part) {
    __shared__ double sm[256];
    double acc = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const double d = dpos[i] >= 0 ? vals[dpos[i]] : 0.0;
        diag[i] = d;
        acc += d;
    }
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sm[threadIdx.x] += sm[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

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

__global__ void k_mp_graph(const int *__restrict__ rp, const int *__restrict__ ci, const double *__restrict__ vals,
                           const double *__restrict__ diag, const double *__restrict__ mean, int n,
                           float *__restrict__ ew, float *__restrict__ cf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double di = diag[i];
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
        const int j = ci[k];
        ew[k] = j == i ? 0.0f : (float)(vals[k] / sqrt(di * diag[j]));
    }
    cf[i] = (float)log(di / mean[0]);
}

// ---- FP32 forward ----------------------------------------------------------------

// sqrt(sum r^2), two stages over the fixed grid
__global__ __launch_bounds__(256) void k_mp_sq_part(const double *__restrict__ r, int n, double *__restrict__ part) {
    __shared__ double sm[256];
    double acc = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) acc += r[i] * r[i];
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sm[threadIdx.x] += sm[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}
__global__ void k_sqrt_inplace(double *__restrict__ v) { v[0] = sqrt(v[0]); }

__global__ __launch_bounds__(256) void k_mp_input(const double *__restrict__ r, const double *__restrict__ rnorm,
                                                  double rsqn, const float *__restrict__ cf,
                                                  const float *__restrict__ w, int n, float *__restrict__ H0) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int i = (int)(e / W), f = (int)(e % W);
    if (i >= n) return;
    const double s = rnorm[0] * rsqn;
    const double inv = s > 0.0 ? 1.0 / s : 0.0;
    H0[e] = (float)(r[i] * inv) * (w[f] + cf[i] * w[W + f]);
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

__global__ __launch_bounds__(256) void k_mp_output(const double *__restrict__ r, const double *__restrict__ rnorm,
                                                   double rsqn, const double *__restrict__ diag,
                                                   const float *__restrict__ wd, const float *__restrict__ H3, int n,
                                                   double *__restrict__ z) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int i = (int)(e / W), f = (int)(e % W);
    float y = i < n ? wd[f] * H3[e] : 0.0f;
    for (int o = W / 2; o > 0; o >>= 1) y += __shfl_xor(y, o, W);
    if (f == 0 && i < n) z[i] = (r[i] + rnorm[0] * rsqn * (double)y) / diag[i];
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

// dy_i = dL/dz_i * s / a_ii; dH3_i = dy_i w_d; gradient of w_d.
__global__ __launch_bounds__(256) void k_mp_out_bwd(const double *__restrict__ r, const double *__restrict__ az,
                                                    const double *__restrict__ rnorm, double rsqn,
                                                    const double *__restrict__ diag, const double *__restrict__ scal,
                                                    int ns, const float *__restrict__ wd, const float *__restrict__ H3,
                                                    int n, float *__restrict__ dH, float *__restrict__ part) {
    __shared__ float sm[kNB][W];
    const int g = threadIdx.x / W, f = threadIdx.x % W;
    const double s = rnorm[0] * rsqn, re = scal[1];
    float acc = 0.0f;
    for (int base = blockIdx.x * kNB; base < n; base += gridDim.x * kNB) {
        const int i = base + g;
        if (i < n) {
            const float dy = (float)(2.0 * (az[i] - r[i]) / re / ns * s / diag[i]);
            const int64_t e = (int64_t)i * W + f;
            dH[e] = dy * wd[f];
            acc += dy * H3[e];
        }
    }
    group_reduce_add(acc, sm, part + (int64_t)blockIdx.x * kMpP + kMpOffWd);
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

__global__ __launch_bounds__(256) void k_mp_in_bwd(const double *__restrict__ r, const double *__restrict__ rnorm,
                                                   double rsqn, const float *__restrict__ cf,
                                                   const float *__restrict__ dH0, int n, float *__restrict__ part) {
    __shared__ float sm[kNB][W];
    const int g = threadIdx.x / W, f = threadIdx.x % W;
    const double s = rnorm[0] * rsqn;
    const double inv = s > 0.0 ? 1.0 / s : 0.0;
    float ae = 0.0f, ac = 0.0f;
    for (int base = blockIdx.x * kNB; base < n; base += gridDim.x * kNB) {
        const int i = base + g;
        if (i < n) {
            const float u = (float)(r[i] * inv), d = dH0[(int64_t)i * W + f];
            ae += u * d;
            ac += u * cf[i] * d;
        }
    }
    group_reduce_add(ae, sm, part + (int64_t)blockIdx.x * kMpP);
    group_reduce_add(ac, sm, part + (int64_t)blockIdx.x * kMpP + W);
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

// BF16 copy: for layer l, output block nb (features 16nb..16nb+15), k chunk kk, lane, i:
// B[k][col] = [W1_l | W2_l][col][k] with k = 16kk + 4(lane/16) + i, col = 16nb + lane%16
// (the B-operand layout of v_mfma_f32_16x16x16_bf16); then w_e, w_c, w_d.
__global__ void k_mp_pack(const float *__restrict__ w, uint16_t *__restrict__ wb) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < 3 * kMpBf16Layer) {
        const int i = idx & 3, lane = (idx >> 2) & 63, kk = (idx >> 8) & 3, nb = (idx >> 10) & 1, l = idx >> 11;
        const int k = 16 * kk + 4 * (lane >> 4) + i, col = 16 * nb + (lane & 15);
        const float *wl = w + kMpOffW + l * 2 * W * W;
        const float v = k < W ? wl[col * W + k] : wl[W * W + col * W + (k - W)];
        wb[idx] = f2bf(v);
    } else if (idx < 3 * kMpBf16Layer + 3 * W) {
        const int t = idx - 3 * kMpBf16Layer;  // 0..95: w_e, w_c, w_d
        const float v = t < 2 * W ? w[t] : w[kMpOffWd + (t - 2 * W)];
        wb[idx] = f2bf(v);
    }
}

// ---- BF16 inference --------------------------------------------------------------

__global__ __launch_bounds__(256) void k_mpb_input(const double *__restrict__ r, const double *__restrict__ rnorm,
                                                   double rsqn, const float *__restrict__ cf,
                                                   const uint16_t *__restrict__ wv, int n, uint16_t *__restrict__ H0,
                                                   const PcgScalars *__restrict__ S) {
    if (gated(S)) return;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int i = (int)(e / W), f = (int)(e % W);
    if (i >= n) return;
    const double s = rnorm[0] * rsqn;
    const double inv = s > 0.0 ? 1.0 / s : 0.0;
    H0[e] = f2bf((float)(r[i] * inv) * (bf2f(wv[f]) + cf[i] * bf2f(wv[W + f])));
}

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

// z_i = (r_i + s y_i) / a_ii, y_i = w_d . h3_i; optional fixed-grid partial sums of r.z and z.ap.
__global__ __launch_bounds__(256) void k_mpb_output(const double *__restrict__ r, const double *__restrict__ rnorm,
                                                    double rsqn, const double *__restrict__ diag,
                                                    const uint16_t *__restrict__ wv, const uint16_t *__restrict__ H3,
                                                    int n, double *__restrict__ z, const double *__restrict__ ap,
                                                    double *__restrict__ rz_part, double *__restrict__ zap_part,
                                                    const PcgScalars *__restrict__ S) {
    if (gated(S)) return;
    __shared__ double sm[2][256];
    const int g = threadIdx.x / W, f = threadIdx.x % W;
    const double s = rnorm[0] * rsqn;
    const float wdf = bf2f(wv[2 * W + f]);
    double arz = 0.0, azap = 0.0;
    for (int base = blockIdx.x * kNB; base < n; base += gridDim.x * kNB) {
        const int i = base + g;
        float y = i < n ? wdf * bf2f(H3[(int64_t)i * W + f]) : 0.0f;
        for (int o = W / 2; o > 0; o >>= 1) y += __shfl_xor(y, o, W);
        if (f == 0 && i < n) {
            const double zi = (r[i] + s * (double)y) / diag[i];
            z[i] = zi;
            arz += r[i] * zi;
            if (ap) azap += zi * ap[i];
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
        if (zap_part) zap_part[blockIdx.x] = sm[1][0];
    }
}

void ensure_ws(Mpnn &net, int n) {
    if (net.ws_n >= n) return;
    const size_t nw = (size_t)std::max(n, 1) * W;
    net.H.alloc(4 * nw);
    net.Mb.alloc(3 * nw);
    net.Pre.alloc(3 * nw);
    net.dH.alloc(nw);
    net.dHp.alloc(nw);
    net.dM.alloc(nw);
    net.hb0.alloc(nw);
    net.hb1.alloc(nw);
    net.z.alloc(std::max(n, 1));
    net.az.alloc(std::max(n, 1));
    net.ws_n = n;
}

void ensure_fixed(Mpnn &net) {
    if (net.part.n) return;
    net.part.alloc((size_t)kMpBlocks * kMpP);
    net.dpart.alloc(3 * kRedBlocks);
    net.scal.alloc(8);
}

}  // namespace

// ---- host side -------------------------------------------------------------------

void mp_graph_prepare(Matrix &A, hipStream_t st) {
    if (A.mp && A.mp->n == A.n) return;
    if (A.diag_status)
        throw Error(MG_E_NOT_PD, "The diagonal element is zero or invalid while building the learned preconditioner.");
    auto g = std::make_shared<MpGraph>();
    const int n = A.n;
    g->n = n;
    g->ew.alloc(std::max<int64_t>(A.nnz_full, 1));
    g->cf.alloc(std::max(n, 1));
    g->diag.alloc(std::max(n, 1));
    DevArray<double> part(kRedBlocks), mean(1);
    hipLaunchKernelGGL(k_mp_diag, dim3(kRedBlocks), dim3(256), 0, st, A.dpos.p, A.vals.p, n, g->diag.p, part.p);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, st, part.p, kRedBlocks, n ? 1.0 / n : 1.0, mean.p);
    if (n > 0)
        hipLaunchKernelGGL(k_mp_graph, dim3(nblk(n, 256)), dim3(256), 0, st, A.rowptr.p, A.colidx.p, A.vals.p,
                           g->diag.p, mean.p, n, g->ew.p, g->cf.p);
    MG_HIP(hipGetLastError());
    MG_HIP(hipStreamSynchronize(st));
    A.mp = g;
}

// Deterministic initialisation: W1, W2 ~ N(0, 1/(2*32)), w_e, w_c ~ N(0, 1), w_d ~ N(0, 0.01^2)
// (a small decoder starts the correction near plain Jacobi).
void mpnn_init(Mpnn &net, uint64_t seed, hipStream_t st) {
    std::mt19937_64 gen(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::vector<float> w(kMpP);
    for (int p = 0; p < kMpP; ++p) {
        double sd = 1.0;
        if (p >= kMpOffW && p < kMpOffWd) sd = 1.0 / std::sqrt(2.0 * W);
        else if (p >= kMpOffWd) sd = 0.01;
        w[p] = (float)(nd(gen) * sd);
    }
    net.w.upload(w);
    std::vector<float> zero(kMpP, 0.0f);
    net.grad.upload(zero);
    net.am.upload(zero);
    net.av.upload(zero);
    net.adam_t = 0;
    net.wb.alloc(3 * kMpBf16Layer + 3 * W);
    mpnn_pack_bf16(net, st);
}

void mpnn_pack_bf16(Mpnn &net, hipStream_t st) {
    const int total = 3 * kMpBf16Layer + 3 * W;
    hipLaunchKernelGGL(k_mp_pack, dim3(nblk(total, 256)), dim3(256), 0, st, net.w.p, net.wb.p);
    MG_HIP(hipGetLastError());
}

namespace {

// forward in FP32 keeping the activations for the backward pass; z into d_z, ||r|| in scal[0]
void forward_f32(Mpnn &net, Matrix &A, const double *d_r, double *d_z, hipStream_t st) {
    const int n = A.n;
    const MpGraph &g = *A.mp;
    const double rsqn = 1.0 / std::sqrt((double)std::max(n, 1));
    const int64_t nw = (int64_t)n * W;
    hipLaunchKernelGGL(k_mp_sq_part, dim3(kRedBlocks), dim3(256), 0, st, d_r, n, net.dpart.p);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, st, net.dpart.p, kRedBlocks, 1.0, net.scal.p);
    hipLaunchKernelGGL(k_sqrt_inplace, dim3(1), dim3(1), 0, st, net.scal.p);
    hipLaunchKernelGGL(k_mp_input, dim3(nblk(nw, 256)), dim3(256), 0, st, d_r, net.scal.p, rsqn, g.cf.p, net.w.p, n,
                       net.H.p);
    for (int l = 0; l < 3; ++l)
        hipLaunchKernelGGL(k_mp_layer, dim3(kMpBlocks), dim3(256), 0, st, net.w.p + kMpOffW + l * 2 * W * W,
                           A.rowptr.p, A.colidx.p, g.ew.p, n, net.H.p + l * nw, net.Mb.p + l * nw, net.Pre.p + l * nw,
                           net.H.p + (l + 1) * nw, l < 2 ? 1 : 0);
    hipLaunchKernelGGL(k_mp_output, dim3(nblk(nw, 256)), dim3(256), 0, st, d_r, net.scal.p, rsqn, g.diag.p,
                       net.w.p + kMpOffWd, net.H.p + 3 * nw, n, d_z);
}

// loss and gradient of one pair, accumulated into net.part and scal[2]
void sample_grad(Mpnn &net, Matrix &A, const double *d_r, const double *d_e, int ns, hipStream_t st) {
    const int n = A.n;
    const MpGraph &g = *A.mp;
    const double rsqn = 1.0 / std::sqrt((double)std::max(n, 1));
    const int64_t nw = (int64_t)n * W;
    forward_f32(net, A, d_r, net.z.p, st);
    spmv_device(A, 1, net.z.p, n, net.az.p, n, st);
    hipLaunchKernelGGL(k_mp_loss_part, dim3(kRedBlocks), dim3(256), 0, st, net.z.p, net.az.p, d_r, d_e, n,
                       net.dpart.p);
    hipLaunchKernelGGL(k_mp_loss_final, dim3(1), dim3(256), 0, st, net.dpart.p, ns, net.scal.p);
    hipLaunchKernelGGL(k_mp_out_bwd, dim3(kMpBlocks), dim3(256), 0, st, d_r, net.az.p, net.scal.p, rsqn, g.diag.p,
                       net.scal.p, ns, net.w.p + kMpOffWd, net.H.p + 3 * nw, n, net.dH.p, net.part.p);
    float *dh = net.dH.p, *dhp = net.dHp.p;
    for (int l = 2; l >= 0; --l) {
        hipLaunchKernelGGL(k_mp_layer_bwd, dim3(kMpBlocks), dim3(256), 0, st, net.w.p + kMpOffW + l * 2 * W * W, n,
                           dh, net.Pre.p + l * nw, net.H.p + l * nw, net.Mb.p + l * nw, l < 2 ? 1 : 0, dhp, net.dM.p,
                           net.part.p + kMpOffW + l * 2 * W * W);
        hipLaunchKernelGGL(k_mp_agg_bwd, dim3(nblk(nw, 256)), dim3(256), 0, st, A.rowptr.p, A.colidx.p, g.ew.p, n,
                           net.dM.p, dhp);
        std::swap(dh, dhp);
    }
    hipLaunchKernelGGL(k_mp_in_bwd, dim3(kMpBlocks), dim3(256), 0, st, d_r, net.scal.p, rsqn, g.cf.p, dh, n,
                       net.part.p);
}

void begin_batch(Mpnn &net, hipStream_t st) {
    const int64_t cnt = (int64_t)kMpBlocks * kMpP;
    hipLaunchKernelGGL(k_fill_f32, dim3(nblk(cnt, 256)), dim3(256), 0, st, net.part.p, cnt, 0.0f);
    hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(3), 0, st, net.scal.p, 3, 0.0);
}

void end_batch(Mpnn &net, hipStream_t st) {
    hipLaunchKernelGGL(k_mp_grad_reduce, dim3(nblk(kMpP, 256)), dim3(256), 0, st, net.part.p, kMpBlocks, net.grad.p);
}

}  // namespace

void mpnn_apply_f32(Mpnn &net, Matrix &A, const double *d_r, double *d_z, hipStream_t st) {
    mp_graph_prepare(A, st);
    ensure_ws(net, A.n);
    ensure_fixed(net);
    if (A.n == 0) return;
    forward_f32(net, A, d_r, d_z, st);
    MG_HIP(hipGetLastError());
}

void mpnn_apply_bf16(Mpnn &net, Matrix &A, const double *d_r, double *d_z, const double *d_rnorm,
                     const PcgScalars *S, const double *d_ap, double *rz_part, double *zap_part, hipStream_t st) {
    mp_graph_prepare(A, st);
    ensure_ws(net, A.n);
    ensure_fixed(net);
    const int n = A.n;
    if (n == 0) return;
    const MpGraph &g = *A.mp;
    const double rsqn = 1.0 / std::sqrt((double)n);
    const int64_t nw = (int64_t)n * W;
    if (!d_rnorm) {
        hipLaunchKernelGGL(k_mp_sq_part, dim3(kRedBlocks), dim3(256), 0, st, d_r, n, net.dpart.p);
        hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, st, net.dpart.p, kRedBlocks, 1.0, net.scal.p);
        hipLaunchKernelGGL(k_sqrt_inplace, dim3(1), dim3(1), 0, st, net.scal.p);
        d_rnorm = net.scal.p;
    }
    const uint16_t *wv = net.wb.p + 3 * kMpBf16Layer;
    hipLaunchKernelGGL(k_mpb_input, dim3(nblk(nw, 256)), dim3(256), 0, st, d_r, d_rnorm, rsqn, g.cf.p, wv, n,
                       net.hb0.p, S);
    uint16_t *hin = net.hb0.p, *hout = net.hb1.p;
    const int tiles = (n + 15) / 16;
    const int grid = std::min(nblk(tiles, 4), 2048);
    for (int l = 0; l < 3; ++l) {
        hipLaunchKernelGGL(k_mpb_layer, dim3(grid), dim3(256), 0, st, net.wb.p + l * kMpBf16Layer, A.rowptr.p,
                           A.colidx.p, g.ew.p, n, hin, hout, l < 2 ? 1 : 0, S);
        std::swap(hin, hout);
    }
    hipLaunchKernelGGL(k_mpb_output, dim3(kRedBlocks), dim3(256), 0, st, d_r, d_rnorm, rsqn, g.diag.p, wv, hin, n,
                       d_z, d_ap, rz_part, zap_part, S);
    MG_HIP(hipGetLastError());
}

double mpnn_loss_grad(Mpnn &net, Matrix &A, int ns, const double *d_R, const double *d_E, int64_t ld,
                      hipStream_t st) {
    mp_graph_prepare(A, st);
    ensure_ws(net, A.n);
    ensure_fixed(net);
    begin_batch(net, st);
    for (int t = 0; t < ns; ++t) sample_grad(net, A, d_R + t * ld, d_E + t * ld, ns, st);
    end_batch(net, st);
    MG_HIP(hipGetLastError());
    double scal[8];
    MG_HIP(hipMemcpyAsync(scal, net.scal.p, sizeof(scal), hipMemcpyDeviceToHost, st));
    MG_HIP(hipStreamSynchronize(st));
    return scal[2];
}

namespace {
__global__ void k_mp_epoch_acc(double *__restrict__ scal, double weight) { scal[3] += scal[2] * weight; }
}  // namespace

// Mini-batch Adam over the ns pairs (columns of R and E), `epochs` passes in a seeded
// shuffled order; losses[e] = mean training loss of epoch e (before each batch's update).
void mpnn_train(Mpnn &net, Matrix &A, int ns, const double *d_R, const double *d_E, int64_t ld, int epochs, int batch,
                float lr, uint64_t seed, double *losses, hipStream_t st) {
    if (ns <= 0 || epochs <= 0) return;
    if (batch <= 0) batch = 16;
    mp_graph_prepare(A, st);
    ensure_ws(net, A.n);
    ensure_fixed(net);
    std::mt19937_64 gen(seed);
    std::vector<int> order(ns);
    for (int t = 0; t < ns; ++t) order[t] = t;
    for (int e = 0; e < epochs; ++e) {
        std::shuffle(order.begin(), order.end(), gen);
        hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(8), 0, st, net.scal.p, 8, 0.0);
        for (int b0 = 0; b0 < ns; b0 += batch) {
            const int nb = std::min(batch, ns - b0);
            begin_batch(net, st);
            for (int t = 0; t < nb; ++t) {
                const int s = order[b0 + t];
                sample_grad(net, A, d_R + s * ld, d_E + s * ld, nb, st);
            }
            end_batch(net, st);
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

}  // namespace mg
