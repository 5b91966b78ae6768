/*
 * magla_testing.h -- test hooks of libmagla.so.  No counterpart in MAGLA: they
 * exist so the parity tests can read device state the C# API never exposes.
 */
#ifndef MAGLA_TESTING_H
#define MAGLA_TESTING_H

#include "magla.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Factor block of one supernode after factorization, as MAGLA stores it:
 * f x w column-major over the supernode's pattern rows (unit diagonal, L
 * below), converted to float64 (f, w = front size and width of the supernode). */
int mg_test_numeric_front(const mg_numeric *num, int32_t supernode, double *out);

/* Runs one 16x16x4 MFMA (MG_FP64 or MG_FP32) whose product is D[i][j] = i + 16 j
 * with the operand lane mapping used by the Schur kernel and stores each
 * accumulator register where the kernel's layout rule places it: out[i*16 + j]
 * must equal i + 16 j (256 values). */
int mg_test_mfma_layout(int32_t precision, double *out);

/* Device schedule summary computed on the host (no GPU needed): out[0..9] =
 * fronts (fused supernode chains), levels, arena elements (factor store,
 * contribution pool and their padding), largest front, widest front, kernel
 * launches per factorization, 64x64 Schur tiles of the lookahead part (the
 * trailing update the next panels do not touch), factor-store elements
 * (sum f w), contribution-pool elements, largest sum of contribution-block
 * elements live at once. */
int mg_test_plan_stats(const mg_symbolic *sym, int64_t *out);

/* Placement of every device front, computed on the host (no GPU needed): for
 * front g, out[6 g .. 6 g + 5] = size f, width w, level, parent front (-1 for a
 * root), factor-store offset, contribution-pool offset (arena elements).
 * out holds 6 x fronts values (fronts = mg_test_plan_stats out[0]). */
int mg_test_plan_layout(const mg_symbolic *sym, int64_t *out);

/* Arena offsets of n front entries (i[k], j[k]), i >= j, of a front of size f and
 * width w placed at factor-store offset off and pool offset coff. */
int mg_test_front_addr(int32_t f, int32_t w, int64_t off, int64_t coff, int32_t n, const int32_t *i,
                       const int32_t *j, int64_t *out);
/* One v_mfma_f32_16x16x16_bf16 with the operand layout of the learned preconditioner's
 * layer kernel: d = a b for 16x16 row-major a, b (rounded to bf16), d in FP32. */
int mg_test_mfma_bf16(const float *a, const float *b, float *d);

/* Sets the Schur tile choice (64 or 128 per front-size bucket, MG_SCHUR_BUCKETS entries)
 * for the handle's precision without timing anything. */
int mg_test_set_schur_tiles(mg_numeric *num, const int32_t *tiles);

/* ILU(0) factors in the matrix's pattern (unit-lower L strictly below the diagonal, U from the
 * diagonal), nnz values in CSR order. */
int mg_test_ilu0_factors(const mg_ilu0 *p, double *out);

/* Launch timing.  mg_profile_enable(1) clears the tallies and brackets every
 * factorization launch with HIP events on its stream; mg_profile_read(kind, ...)
 * returns 1 and the kernel's name, launch count, summed event time (ms) and
 * summed algorithmic flops / HBM bytes (kinds 0..6: assemble_pull,
 * assemble_values, panel_diag, panel_rows, schur, forward, backward), 0 past
 * the last kind, or a negative status. */
void mg_profile_enable(int32_t on);
int mg_profile_read(int32_t kind, const char **name, int64_t *launches, double *ms, double *flops, double *bytes);

#ifdef __cplusplus
}
#endif
#endif
