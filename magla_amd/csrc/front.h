// front.h -- where the entries of a device front live.
//
// A front of f rows and w pivot columns is split between the two parts of the numeric arena:
//   * its pivot columns j < w form an f x w column-major block (ld = f) in the factor store,
//     kept after the factorization for the solves;
//   * its trailing t x t block (t = f - w, the contribution read by the parent front) sits in a
//     contribution pool only while it is live.  It keeps the lower trapezoid of each 64-column
//     block: column c of the block (0-based within it) holds rows [64 b, t) of column block b,
//     so the block takes about t^2 / 2 elements instead of t^2.
// Entry (i, j) with i >= j, both front-local, is arena[front_col(off, coff, f, w, j) + i].
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mg {

// Elements of a contribution block's column blocks before column block cb (64 columns each,
// column block c holding rows [64 c, t)).
__host__ __device__ inline int64_t contrib_before(int t, int cb) {
    return (int64_t)64 * cb * t - (int64_t)2048 * cb * (cb - 1);
}

// Elements of a t x t contribution block.
__host__ __device__ inline int64_t contrib_size(int t) {
    if (t <= 0) return 0;
    const int last = ((t - 1) >> 6) << 6;  // first column of the last column block
    return contrib_before(t, last >> 6) + (int64_t)(t - last) * (t - last);
}

// Entry (r, c), r >= c, of a contribution block at pool offset coff is
// arena[contrib_col(coff, t, c) + r].
__host__ __device__ inline int64_t contrib_col(int64_t coff, int t, int c) {
    const int r0 = c & ~63;
    return coff + contrib_before(t, c >> 6) + (int64_t)(c - r0) * (t - r0) - r0;
}

__host__ __device__ inline int64_t front_col(int64_t off, int64_t coff, int f, int w, int j) {
    return j < w ? off + (int64_t)j * f : contrib_col(coff, f - w, j - w) - w;
}

// Columns [jg, jend) of a front as one affine run: entry (i, j) at base + (j - jg) ld + i.
struct ColRun {
    int64_t base;
    int ld, jg, jend;
    __device__ __forceinline__ int64_t at(int i, int j) const { return base + (int64_t)((j - jg) * ld + i); }
};

// A front's placement, as the tile kernels take it.
struct FrontCols {
    int64_t off, coff;
    int f, w;
    __device__ __forceinline__ int64_t col(int j) const { return front_col(off, coff, f, w, j); }
    // Columns [jg, jg + span) of a tile whose first column is j0.  A tile that starts in the
    // pivot columns keeps to them (the run ends at w, and is empty when jg >= w); a tile that
    // starts in the contribution starts at w + 64 b + o with o + span <= 64, so the run lies
    // in one column block (it ends at f).  Tiles are laid out so that this holds for every
    // wave's columns.
    __device__ __forceinline__ ColRun run(int j0, int jg, int span) const {
        if (j0 < w) return ColRun{off + (int64_t)jg * f, f, jg, min(jg + span, w)};
        const int c = jg - w;
        return ColRun{contrib_col(coff, f - w, c) - w, f - w - (c & ~63), jg, min(jg + span, f)};
    }
    // col(j) for a front whose contribution has at most 64 columns (one column block).
    __device__ __forceinline__ int64_t col_small(int j) const {
        return j < w ? off + (int64_t)j * f : coff + (int64_t)((j - w) * (f - w) - w);
    }
    // A safe address for reads whose value is not used: the tile's first diagonal entry.
    __device__ __forceinline__ int64_t diag(int j0) const { return col(j0) + j0; }
};

}  // namespace mg
