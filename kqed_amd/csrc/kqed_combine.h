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
/* X, Y [E][c][4]: the E derived argument pairs of each of the c pairs (x [c][4]; y with stride ys
 * doubles per pair, 0 = one y for all) */
void kqed_comb_launch_args(int comb, size_t c, const double *x, const double *y, int ys, double *X, double *Y, hipStream_t s);
/* From the L0 tensors L [E][c][384]: out [c][384 * kqed_comb_kernels(comb)], or with a site tensor
 * Tc [c][384] (FP32) the per-pair dots out[q * dstride + p] = sum_c K_q(x_p, y_p)_c Tc[p][c] */
void kqed_comb_launch(int comb, size_t c, const double *L, const double *x, const double *y, int ys, const double *M,
                      double *out, const float *Tc, size_t dstride, hipStream_t s);

#endif
