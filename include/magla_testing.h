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

/* Front of one supernode after factorization: f x f column-major (lower part
 * valid), converted to float64.  Columns [0, w) are the factor block (unit
 * diagonal, L below), the trailing (f-w)^2 block is the contribution. */
int mg_test_numeric_front(const mg_numeric *num, int32_t supernode, double *out);

/* Runs one 16x16x4 MFMA (MG_FP64 or MG_FP32) whose product is D[i][j] = i + 16 j
 * with the operand lane mapping used by the Schur kernel and stores each
 * accumulator register where the kernel's layout rule places it: out[i*16 + j]
 * must equal i + 16 j (256 values). */
int mg_test_mfma_layout(int32_t precision, double *out);

#ifdef __cplusplus
}
#endif
#endif
