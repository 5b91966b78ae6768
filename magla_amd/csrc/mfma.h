// mfma.h -- 64x64 Schur-update tile on CDNA4 matrix cores.
//
// C[i0+r, j0+c] -= sum_k As[k][r] * Bs[k][c] for the lower triangle
// (i0+r >= j0+c) of a column-major front with leading dimension f; K is
// accumulated in registers over LDS chunks (schur_acc_tile), then stored once.
//
// Four waves; wave v owns the 32x32 quadrant (rows 32*(v&1), cols 32*(v>>1))
// as 2x2 blocks of 16x16 MFMA tiles, k advancing by 4.  The product is formed
// transposed (A operand = column panel, B operand = row panel) so that
// consecutive lanes own consecutive ROWS of one front column: loads and
// stores of C are row runs.  Accumulator layouts of the 16x16x4 instructions
// (lane l, register q, column l%16): f32 holds row 4*(l/16)+q, f64 holds row
// (l/16)+4q; MfmaAcc<T>::row gives the row for the element type, and the
// layout is checked on the device by mg_test_mfma_layout.
#pragma once

#include <hip/hip_runtime.h>

namespace mg {

typedef double mg_d4 __attribute__((ext_vector_type(4)));
typedef float mg_f4 __attribute__((ext_vector_type(4)));

template <typename T>
struct MfmaAcc;
template <>
struct MfmaAcc<double> {
    typedef mg_d4 type;
    static __device__ inline type step(double a, double b, type c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lk, int q) { return lk + 4 * q; }
};
template <>
struct MfmaAcc<float> {
    typedef mg_f4 type;
    static __device__ inline type step(float a, float b, type c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lk, int q) { return 4 * lk + q; }
};

template <typename T>
__device__ inline void schur_acc_zero(typename MfmaAcc<T>::type (&acc)[2][2]) {
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int q = 0; q < 4; ++q) acc[a][b][q] = T(0);
}

// acc += As^T-panel x Bs-panel over nk (multiple of 4) rows of the LDS chunks.
template <typename T>
__device__ inline void schur_acc_tile(const T (*As)[65], const T (*Bs)[65], int nk,
                                      typename MfmaAcc<T>::type (&acc)[2][2]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rq = (wave & 1) * 32, cq = (wave >> 1) * 32;
    const int l16 = lane & 15, lk = lane >> 4;
    for (int k0 = 0; k0 < nk; k0 += 4) {
        const int k = k0 + lk;
        T colv[2], rowv[2];
        for (int b = 0; b < 2; ++b) colv[b] = Bs[k][cq + 16 * b + l16];
        for (int a = 0; a < 2; ++a) rowv[a] = As[k][rq + 16 * a + l16];
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) acc[a][b] = MfmaAcc<T>::step(colv[b], rowv[a], acc[a][b]);
    }
}

// M[j*f + i] -= C[i][j] for the tile's lower entries with i < f and j < jend, where
// acc[a][b][q] = C[row rq+16a+l16][col cq+16b+MfmaAcc<T>::row(lk,q)].
template <typename T>
__device__ inline void schur_store_tile(const typename MfmaAcc<T>::type (&acc)[2][2], T *__restrict__ M, int f,
                                        int i0, int j0, int jend) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rq = (wave & 1) * 32, cq = (wave >> 1) * 32;
    const int l16 = lane & 15, lk = lane >> 4;
    for (int a = 0; a < 2; ++a) {
        const int i = i0 + rq + 16 * a + l16;
        if (i >= f) continue;
        for (int b = 0; b < 2; ++b)
            for (int q = 0; q < 4; ++q) {
                const int j = j0 + cq + 16 * b + MfmaAcc<T>::row(lk, q);
                if (j <= i && j < jend) M[(int64_t)j * f + i] -= acc[a][b][q];
            }
    }
}

}  // namespace mg
