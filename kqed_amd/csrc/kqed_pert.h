/*
 * kqed_pert.h -- launchers of the perturbative lepton-loop kernels (kqed_pert.hip) for the host API.
 */
#ifndef KQED_PERT_H
#define KQED_PERT_H

#include <hip/hip_runtime.h>
#include <stddef.h>

/* bytes of the lepton loop's gather lists, and their construction from the reference's G8 table
 * (returns -1 when a trace is not 0 or +-4 or a thread has more nonzero terms than the lists hold) */
size_t kqed_pert_table_bytes(void);
int kqed_pert_build_table(const int *G8, unsigned *tab);
/* pihat1 at n pairs: out[n][1280] = v1[4][4][4][4][4], r1[4][4][4][4] per pair */
void kqed_pert_launch_pihat1(size_t n, const double *x, const double *y, const unsigned *tab, double *out, hipStream_t s);
/* ipihatFermLoop_antisym at n pairs: out[n][384] */
void kqed_pert_launch_ipihat(size_t n, const double *x, const double *y, const unsigned *tab, double *out, hipStream_t s);
/* xs = sc x, ys = sc y for n pairs */
void kqed_pert_launch_scale(size_t n, double sc, const double *x, const double *y, double *xs, double *ys, hipStream_t s);

#endif
