/*
 * kqed_pert.h -- launchers of the perturbative lepton-loop kernels (kqed_pert.hip) for the host API.
 */
#ifndef KQED_PERT_H
#define KQED_PERT_H

#include <hip/hip_runtime.h>
#include <stddef.h>

/* bytes of the int8 trace tables, and their construction from the reference's G8 table */
size_t kqed_pert_table_bytes(void);
void kqed_pert_build_table(const int *G8, signed char *tab);
/* pihat1 at n pairs: out[n][1280] = v1[4][4][4][4][4], r1[4][4][4][4] per pair */
void kqed_pert_launch_pihat1(size_t n, const double *x, const double *y, const signed char *tab, double *out, hipStream_t s);
/* ipihatFermLoop_antisym at n pairs: out[n][384]; scratch holds 3 n * 1280 doubles */
void kqed_pert_launch_ipihat(size_t n, const double *x, const double *y, const signed char *tab, double *scratch,
                             double *out, hipStream_t s);
/* xs = sc x, ys = sc y for n pairs */
void kqed_pert_launch_scale(size_t n, double sc, const double *x, const double *y, double *xs, double *ys, hipStream_t s);

#endif
