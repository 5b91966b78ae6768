// mfma.h -- 64x64 Schur-update tile on CDNA4 matrix cores.
//
// C[i0+r, j0+c] -= sum_k As[k][r] * Bs[k][c] for the lower triangle
// (i0+r >= j0+c) of a front whose columns front.h places; K is
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

#include "front.h"

namespace mg {

// Row stride (elements) of the LDS operand chunks.  FP32: 64, with the 16-column groups of row
// k stored XOR-swizzled by (k mod 4) (sch_col): the four k-rows one operand read touches
// (16 lanes of 4 bytes each) then fall on four different 16-bank groups, so the read is free
// of bank conflicts with 64 banks (and with 32 banks served per half wave); a swizzled 4-float
// vector stays contiguous and 16-byte aligned for the stores that fill a row.  FP64: 65 (a
// row of 16 lanes of 8 bytes already spans all banks; four rows take four cycles at any
// stride), no swizzle.
template <typename T>
constexpr int kSchurLd = sizeof(T) == 4 ? 64 : 65;
template <typename T>
__device__ __forceinline__ int sch_col(int k, int c) {
    if constexpr (sizeof(T) == 4) return c ^ ((k & 3) << 4);
    else return c;
}

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

// acc += As^T-panel x Bs-panel over the 64 rows of the LDS chunks (rows past the
// chunk's length hold zeros).  The k-steps go in batches whose operands are read
// from LDS while the previous batch runs on the matrix cores (software pipeline;
// the scheduling barriers keep the compiler from sinking the reads to their uses).
template <typename T>
__device__ inline void schur_read_ops(const T (*As)[kSchurLd<T>], const T (*Bs)[kSchurLd<T>], int k0,
                                      T (&colv)[2], T (&rowv)[2]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rq = (wave & 1) * 32, cq = (wave >> 1) * 32;
    const int k = k0 + (lane >> 4), l16 = lane & 15;
#pragma unroll
    for (int b = 0; b < 2; ++b) colv[b] = Bs[k][sch_col<T>(k, cq + 16 * b + l16)];
#pragma unroll
    for (int a = 0; a < 2; ++a) rowv[a] = As[k][sch_col<T>(k, rq + 16 * a + l16)];
}

template <typename T>
__device__ inline void schur_acc_tile(const T (*As)[kSchurLd<T>], const T (*Bs)[kSchurLd<T>],
                                      typename MfmaAcc<T>::type (&acc)[2][2]) {
    constexpr int kSteps = 4;                 // k-steps (of 4) per batch
    constexpr int kBatches = 64 / (4 * kSteps);
    T colv[2][kSteps][2], rowv[2][kSteps][2];  // double-buffered operand batches
#pragma unroll
    for (int s = 0; s < kSteps; ++s) schur_read_ops<T>(As, Bs, 4 * s, colv[0][s], rowv[0][s]);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int bt = 0; bt < kBatches; ++bt) {
        const int cur = bt & 1;
        if (bt + 1 < kBatches) {
#pragma unroll
            for (int s = 0; s < kSteps; ++s)
                schur_read_ops<T>(As, Bs, 4 * (kSteps * (bt + 1) + s), colv[cur ^ 1][s], rowv[cur ^ 1][s]);
        }
#pragma unroll
        for (int s = 0; s < kSteps; ++s)
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    acc[a][b] = MfmaAcc<T>::step(colv[cur][s][b], rowv[cur][s][a], acc[a][b]);
        __builtin_amdgcn_sched_barrier(0);
    }
}

// F[i][j] -= C[i][j] for the tile's lower entries with i < f and j < jend, where
// acc[a][b][q] = C[row rq+16a+l16][col cq+16b+MfmaAcc<T>::row(lk,q)].  schur_load_tile issues
// the 16 reads of the old values together (before the last chunk's products, so that their
// latency overlaps the matrix cores' work); schur_store_tile subtracts on every lane (which
// keeps the reads unconditional, where they were issued) and writes the tile's entries.
template <typename T>
__device__ inline void schur_load_tile(T (&old)[16], const T *__restrict__ arena, const FrontCols &fr, int i0,
                                       int j0) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rq = (wave & 1) * 32, cq = (wave >> 1) * 32;
    const int l16 = lane & 15, lk = lane >> 4;
    const int f = fr.f;
    const ColRun cr = fr.run(j0, j0 + cq, 32);
    const int64_t fb = fr.diag(j0);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = i0 + rq + 16 * a + l16;
                const int j = j0 + cq + 16 * b + MfmaAcc<T>::row(lk, q);
                const int ic = min(i, f - 1);  // unconditional load: entries not stored read fb
                old[(2 * a + b) * 4 + q] = arena[(j < cr.jend && j <= ic) ? cr.at(ic, j) : fb];
            }
}

template <typename T>
__device__ inline void schur_store_tile(const typename MfmaAcc<T>::type (&acc)[2][2], const T (&old)[16],
                                        T *__restrict__ arena, const FrontCols &fr, int i0, int j0, int jend) {
    const int f = fr.f;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rq = (wave & 1) * 32, cq = (wave >> 1) * 32;
    const int l16 = lane & 15, lk = lane >> 4;
    const ColRun cr = fr.run(j0, j0 + cq, 32);
    jend = min(jend, cr.jend);
    T v[16];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) v[(2 * a + b) * 4 + q] = old[(2 * a + b) * 4 + q] - acc[a][b][q];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = i0 + rq + 16 * a + l16;
                const int j = j0 + cq + 16 * b + MfmaAcc<T>::row(lk, q);
                if (i < f && j <= i && j < jend) arena[cr.at(i, j)] = v[(2 * a + b) * 4 + q];
            }
}

// ---- 128x128 tiles (large fronts) --------------------------------------------------
//
// Four waves as 2x2; wave v owns the 64x64 quadrant (rows 64*(v&1), cols 64*(v>>1)) as
// 4x4 blocks of 16x16 MFMA tiles, so every operand read from LDS feeds four MFMAs and a
// workgroup does twice the products of a 64x64 tile per byte it loads.  LDS chunks hold
// kT128K<T> k-rows of 128 (+16 pad, which puts rows k and k+1 in disjoint bank halves).
constexpr int kT128Pad = 128 + 16;
template <typename T>
constexpr int kT128K = sizeof(T) == 4 ? 32 : 16;  // two buffers of both operands fit twice per CU

template <typename T>
__device__ inline void schur128_acc_zero(typename MfmaAcc<T>::type (&acc)[4][4]) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[a][b][q] = T(0);
}

template <typename T>
__device__ inline void schur128_read_ops(const T (*As)[kT128Pad], const T (*Bs)[kT128Pad], int k0, T (&colv)[4],
                                         T (&rowv)[4]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rq = (wave & 1) * 64, cq = (wave >> 1) * 64;
    const int k = k0 + (lane >> 4), l16 = lane & 15;
#pragma unroll
    for (int b = 0; b < 4; ++b) colv[b] = Bs[k][cq + 16 * b + l16];
#pragma unroll
    for (int a = 0; a < 4; ++a) rowv[a] = As[k][rq + 16 * a + l16];
}

template <typename T>
__device__ inline void schur128_step(const T (&colv)[4], const T (&rowv)[4], typename MfmaAcc<T>::type (&acc)[4][4]) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = MfmaAcc<T>::step(colv[b], rowv[a], acc[a][b]);
}

// The old values of the wave's outputs in row blocks [a0, a0 + NA): unconditional reads (an
// entry that is not stored reads the tile's first diagonal entry), issued together.  old[(a - a0) * 16 + b * 4 + q] holds output
// (row rq+16a+l16, col cq+16b+MfmaAcc<T>::row(lk,q)).
template <typename T, int NA, int a0>
__device__ inline void schur128_load_old(T (&old)[NA * 16], const T *__restrict__ arena, const FrontCols &fr, int i0,
                                         int j0) {
    const int f = fr.f;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rq = (wave & 1) * 64, cq = (wave >> 1) * 64;
    const int l16 = lane & 15, lk = lane >> 4;
    const ColRun cr = fr.run(j0, j0 + cq, 64);
    const int64_t fb = fr.diag(j0);
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        const int ic = min(i0 + rq + 16 * (a0 + a) + l16, f - 1);
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int j = j0 + cq + 16 * b + MfmaAcc<T>::row(lk, q);
                old[a * 16 + b * 4 + q] = arena[(j < cr.jend && j <= ic) ? cr.at(ic, j) : fb];
            }
    }
}

// F[i][j] = old - C[i][j] on the lower entries with i < f, j < jend of row blocks
// [a0, a0 + NA); the differences are formed on every lane (the reads stay where issued).
template <typename T, int NA, int a0>
__device__ inline void schur128_store_old(const typename MfmaAcc<T>::type (&acc)[4][4], const T (&old)[NA * 16],
                                          T *__restrict__ arena, const FrontCols &fr, int i0, int j0, int jend) {
    const int f = fr.f;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rq = (wave & 1) * 64, cq = (wave >> 1) * 64;
    const int l16 = lane & 15, lk = lane >> 4;
    const ColRun cr = fr.run(j0, j0 + cq, 64);
    jend = min(jend, cr.jend);
    T v[NA * 16];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) v[a * 16 + b * 4 + q] = old[a * 16 + b * 4 + q] - acc[a0 + a][b][q];
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        const int i = i0 + rq + 16 * (a0 + a) + l16;
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int j = j0 + cq + 16 * b + MfmaAcc<T>::row(lk, q);
                if (i < f && j <= i && j < jend) arena[cr.at(i, j)] = v[a * 16 + b * 4 + q];
            }
    }
}

}  // namespace mg
