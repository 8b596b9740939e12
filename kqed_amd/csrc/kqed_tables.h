/*
 * kqed_tables.h -- host-side table subsystem: the three checksummed table files of the
 * reference (form factors, Taylor-x, Taylor-y), their CRC-32C checksum scheme, and a seeded
 * generator for a synthetic form-factor table in the reference's exact file format.
 */
#ifndef KQED_TABLES_H
#define KQED_TABLES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KQED_MAGIC 816968
#define KQED_NTAYX_ROWS 23   /* rows of taylorx_cksum.bin, reference KQED.h:57 */
#define KQED_NTAYX_COLS 810
#define KQED_NTAYY_ROWS 14   /* rows of taylory_cksum.bin, reference KQED.h:60 */
#define KQED_NTAYY_COLS 100
#define KQED_NFFA 14         /* form-factor tables, reference KQED.h:52-54 */

/* Host copy of every table.  Form factors are flattened [nm][ix][iy][k] with k < nfmax and
 * zero padding beyond nfx[ix] (the reference keeps ragged float**** rows, io.c:106-158). */
typedef struct {
  int nstpx, nstpy, nfmax;
  double *XX, *YY;
  int *nfx;
  float *Ffm, *Ffp;
  double TX[KQED_NTAYX_ROWS][KQED_NTAYX_COLS];
  double TY[KQED_NTAYY_ROWS][KQED_NTAYY_COLS];
} kqed_host_tables;

uint32_t kqed_crc32c(uint32_t crc, const void *buf, size_t len);
/* DML_checksum_accum_crc32c of the reference (crc32c.c:160-172) */
void kqed_cksum_accum(uint32_t *a, uint32_t *b, uint32_t rank, const void *buf, size_t len);

/* Readers (reference io.c:42-288).  Return 0 on success, non-zero on error with a message
 * on stderr: 1 missing file, 2 bad magic, 3 bad dimension, 4 checksum mismatch, 5 short read. */
int kqed_read_ff(const char *path, kqed_host_tables *t);
int kqed_read_taylorx(const char *path, kqed_host_tables *t);
int kqed_read_taylory(const char *path, kqed_host_tables *t);
void kqed_free_host_tables(kqed_host_tables *t);

/* Writers in the same format (reference io.c:292-413). */
int kqed_write_ff(const char *path, const kqed_host_tables *t);
int kqed_write_taylor(const char *path, const double *rows, int nrow, int ncol);

/* Synthetic form-factor table.  Grid: XX[i] = xh*k_i with k_i = 3(i+1) for i <= 14 and
 * 45 + 9(i-14) beyond (i < nstpx), YY[i] = yh*(i+1) (i < nstpy),
 * nfx[ix] = nf0 + (nf1 - nf0) * ix / nstpx.  Coefficients follow the Chebyshev-U expansion
 * structure of the weight functions: c_j(x,y) = A(x,y) w_j rho^j, rho = min(x,y)/max(x,y), stored
 * through the reference's y-power convention so that the interpolator rebuilds c_j; the d/dx and
 * d2/dx2 tables hold the analytic x-derivatives.  Deterministic in seed. */
typedef struct {
  int nstpx, nstpy, nf0, nf1;
  double xh, yh;
  uint64_t seed;
} kqed_synth_params;
int kqed_synth_ff(const char *path, const kqed_synth_params *p);

#ifdef __cplusplus
}
#endif
#endif
