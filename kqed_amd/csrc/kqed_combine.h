/*
 * kqed_combine.h -- launchers of the kernel-combination kernels (kqed_combine.hip) for the host
 * API: derived argument pairs, then the per-pair combination of their L0 tensors.
 */
#ifndef KQED_COMBINE_H
#define KQED_COMBINE_H

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "../../include/kqed.h"

/* L0 evaluations per pair of combination `comb` (0 for an unknown comb) */
int kqed_comb_evals(int comb);
/* X, Y [E][c][4]: the E derived argument pairs of each of the c pairs (x, y [c][4]) */
void kqed_comb_launch_args(int comb, size_t c, const double *x, const double *y, double *X, double *Y, hipStream_t s);
/* out [c][384 * kqed_comb_kernels(comb)] from the L0 tensors L [E][c][384] */
void kqed_comb_launch(int comb, size_t c, const double *L, const double *x, const double *y, const double *M, double *out,
                      hipStream_t s);

#endif
