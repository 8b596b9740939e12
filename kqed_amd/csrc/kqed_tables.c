/*
 * kqed_tables.c -- table files: reader, writer, checksums and the synthetic form-factor table.
 *
 * File formats follow the reference loader (io.c:42-288) byte for byte, little-endian:
 *   FF : int magic, int nstpx, double XX[nstpx], u32 ck[2], int nstpy, double YY[nstpy], u32 ck[2],
 *        int Nffa, then per table {int nstpx, per x node {int nstpy, per y node {int nx,
 *        float Fm[nx], float Fp[nx]}}}, u32 ck[4]   (checksum rank = running row counter)
 *   TX : int magic, int 23, 23 x {int 810, double[810]}, u32 ck[2]   (rank = row)
 *   TY : int magic, int 14, 14 x {int 100, double[100]}, u32 ck[2]
 */
#include "kqed_tables.h"
#include "../../include/kqed.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

/* ------------------------------------------------------------------ CRC-32C */
static uint32_t g_crc[256];
static int g_crc_ok = 0;

uint32_t kqed_crc32c(uint32_t crc, const void *buf, size_t len)
{
  if (!g_crc_ok) {
    for (uint32_t i = 0; i < 256; i++) {
      uint32_t c = i;
      for (int k = 0; k < 8; k++) c = (c >> 1) ^ (0x82F63B78u & (0u - (c & 1u)));
      g_crc[i] = c;
    }
    g_crc_ok = 1;
  }
  const uint8_t *p = (const uint8_t *)buf;
  size_t n = len & ~(size_t)3; /* whole 32-bit words, as the reference's SSE4.2 path */
  while (n--) crc = g_crc[(crc ^ *p++) & 0xFFu] ^ (crc >> 8);
  return crc;
}

static inline uint32_t rot(uint32_t w, uint32_t r) { return r ? (w << r) | (w >> (32u - r)) : w; }

void kqed_cksum_accum(uint32_t *a, uint32_t *b, uint32_t rank, const void *buf, size_t len)
{
  const uint32_t w = kqed_crc32c(0, buf, len);
  *a ^= rot(w, rank % 29u);
  *b ^= rot(w, rank % 31u);
}

/* ------------------------------------------------------------------ reading */
typedef struct { const uint8_t *p, *end; } cursor;

static int take(cursor *c, void *dst, size_t n)
{
  if ((size_t)(c->end - c->p) < n) return 5;
  memcpy(dst, c->p, n);
  c->p += n;
  return 0;
}

static uint8_t *slurp(const char *path, size_t *len)
{
  FILE *f = fopen(path, "rb");
  if (!f) return NULL;
  struct stat st;
  if (fstat(fileno(f), &st) != 0) { fclose(f); return NULL; }
  uint8_t *b = (uint8_t *)malloc((size_t)st.st_size + 1);
  if (!b) { fclose(f); return NULL; }
  *len = fread(b, 1, (size_t)st.st_size, f);
  fclose(f);
  return b;
}

int kqed_read_ff(const char *path, kqed_host_tables *t)
{
  size_t len = 0;
  uint8_t *buf = slurp(path, &len);
  if (!buf) { fprintf(stderr, "kqed: cannot open form-factor table %s\n", path); return 1; }
  cursor c = { buf, buf + len };
  int rc = 0, magic = 0, n = 0, nffa = 0;
  uint32_t ck[4], a, b;
  if ((rc = take(&c, &magic, 4))) goto out;
  if (magic != KQED_MAGIC) { rc = 2; fprintf(stderr, "kqed: %s: missing magic number\n", path); goto out; }
  if ((rc = take(&c, &t->nstpx, 4))) goto out;
  if (t->nstpx < 2 || t->nstpx > KQED_NTAYY_COLS) {
    /* Tabd_yeq0 indexes the Taylor-y rows with the x-node index (reference Tabd.c:127-188) */
    rc = 3; fprintf(stderr, "kqed: %s: nstpx=%d outside [2,%d]\n", path, t->nstpx, KQED_NTAYY_COLS); goto out;
  }
  t->XX = (double *)malloc(sizeof(double) * t->nstpx);
  if ((rc = take(&c, t->XX, 8u * t->nstpx)) || (rc = take(&c, ck, 8))) goto out;
  a = b = 0; kqed_cksum_accum(&a, &b, 0, t->XX, 8u * t->nstpx);
  if (a != ck[0] || b != ck[1]) { rc = 4; fprintf(stderr, "kqed: %s: XX checksum mismatch\n", path); goto out; }
  if ((rc = take(&c, &t->nstpy, 4))) goto out;
  if (t->nstpy < 2) { rc = 3; fprintf(stderr, "kqed: %s: nstpy=%d\n", path, t->nstpy); goto out; }
  t->YY = (double *)malloc(sizeof(double) * t->nstpy);
  if ((rc = take(&c, t->YY, 8u * t->nstpy)) || (rc = take(&c, ck, 8))) goto out;
  a = b = 0; kqed_cksum_accum(&a, &b, 0, t->YY, 8u * t->nstpy);
  if (a != ck[0] || b != ck[1]) { rc = 4; fprintf(stderr, "kqed: %s: YY checksum mismatch\n", path); goto out; }
  if ((rc = take(&c, &nffa, 4))) goto out;
  if (nffa != KQED_NFFA) { rc = 3; fprintf(stderr, "kqed: %s: Nffa=%d, expected 14\n", path, nffa); goto out; }
  /* pass 1: row lengths; one length per x node (as the reference's GPU loader requires) */
  t->nfx = (int *)calloc(t->nstpx, sizeof(int));
  t->nfmax = 0;
  {
    cursor c1 = c;
    for (int nm = 0; nm < nffa; nm++) {
      if ((rc = take(&c1, &n, 4))) goto out;
      if (n != t->nstpx) { rc = 3; fprintf(stderr, "kqed: %s: table %d has %d x nodes\n", path, nm, n); goto out; }
      for (int ix = 0; ix < t->nstpx; ix++) {
        if ((rc = take(&c1, &n, 4))) goto out;
        if (n != t->nstpy) { rc = 3; fprintf(stderr, "kqed: %s: table %d x node %d has %d y nodes\n", path, nm, ix, n); goto out; }
        for (int iy = 0; iy < t->nstpy; iy++) {
          int nx;
          if ((rc = take(&c1, &nx, 4))) goto out;
          if (nx < 1 || nx > 4096 || (nm + iy > 0 && nx != t->nfx[ix])) {
            rc = 3; fprintf(stderr, "kqed: %s: row length %d at (%d,%d,%d)\n", path, nx, nm, ix, iy); goto out;
          }
          t->nfx[ix] = nx;
          if (nx > t->nfmax) t->nfmax = nx;
          if ((size_t)(c1.end - c1.p) < 8u * (size_t)nx) { rc = 5; goto out; }
          c1.p += 8u * (size_t)nx;
        }
      }
    }
  }
  {
    const size_t tot = (size_t)KQED_NFFA * t->nstpx * t->nstpy * t->nfmax;
    t->Ffm = (float *)calloc(tot, sizeof(float));
    t->Ffp = (float *)calloc(tot, sizeof(float));
    if (!t->Ffm || !t->Ffp) { rc = 5; fprintf(stderr, "kqed: out of host memory\n"); goto out; }
    uint32_t ma = 0, mb = 0, pa = 0, pb = 0, rank = 0;
    for (int nm = 0; nm < nffa; nm++) {
      c.p += 4;
      for (int ix = 0; ix < t->nstpx; ix++) {
        c.p += 4;
        for (int iy = 0; iy < t->nstpy; iy++) {
          int nx;
          take(&c, &nx, 4);
          const size_t o = (((size_t)nm * t->nstpx + ix) * t->nstpy + iy) * t->nfmax;
          take(&c, t->Ffm + o, 4u * nx);
          kqed_cksum_accum(&ma, &mb, rank, t->Ffm + o, 4u * nx);
          take(&c, t->Ffp + o, 4u * nx);
          kqed_cksum_accum(&pa, &pb, rank, t->Ffp + o, 4u * nx);
          rank++;
        }
      }
    }
    if ((rc = take(&c, ck, 16))) goto out;
    if (ck[0] != ma || ck[1] != mb || ck[2] != pa || ck[3] != pb) {
      rc = 4; fprintf(stderr, "kqed: %s: form-factor checksum mismatch\n", path); goto out;
    }
  }
out:
  if (rc == 5) fprintf(stderr, "kqed: %s: file is truncated\n", path);
  free(buf);
  return rc;
}

static int read_taylor(const char *path, double *rows, int nrow, int ncol)
{
  size_t len = 0;
  uint8_t *buf = slurp(path, &len);
  if (!buf) { fprintf(stderr, "kqed: cannot open Taylor table %s\n", path); return 1; }
  cursor c = { buf, buf + len };
  int rc = 0, magic, n;
  uint32_t a = 0, b = 0, ck[2];
  if ((rc = take(&c, &magic, 4))) goto out;
  if (magic != KQED_MAGIC) { rc = 2; fprintf(stderr, "kqed: %s: missing magic number\n", path); goto out; }
  if ((rc = take(&c, &n, 4))) goto out;
  if (n != nrow) { rc = 3; fprintf(stderr, "kqed: %s: %d rows, expected %d\n", path, n, nrow); goto out; }
  for (int i = 0; i < nrow; i++) {
    if ((rc = take(&c, &n, 4))) goto out;
    if (n != ncol) { rc = 3; fprintf(stderr, "kqed: %s: row %d has %d entries, expected %d\n", path, i, n, ncol); goto out; }
    if ((rc = take(&c, rows + (size_t)i * ncol, 8u * ncol))) goto out;
    kqed_cksum_accum(&a, &b, (uint32_t)i, rows + (size_t)i * ncol, 8u * ncol);
  }
  if ((rc = take(&c, ck, 8))) goto out;
  if (ck[0] != a || ck[1] != b) { rc = 4; fprintf(stderr, "kqed: %s: checksum mismatch\n", path); goto out; }
out:
  if (rc == 5) fprintf(stderr, "kqed: %s: file is truncated\n", path);
  free(buf);
  return rc;
}

int kqed_read_taylorx(const char *path, kqed_host_tables *t)
{ return read_taylor(path, &t->TX[0][0], KQED_NTAYX_ROWS, KQED_NTAYX_COLS); }
int kqed_read_taylory(const char *path, kqed_host_tables *t)
{ return read_taylor(path, &t->TY[0][0], KQED_NTAYY_ROWS, KQED_NTAYY_COLS); }

void kqed_free_host_tables(kqed_host_tables *t)
{
  free(t->XX); free(t->YY); free(t->nfx); free(t->Ffm); free(t->Ffp);
  t->XX = t->YY = NULL; t->nfx = NULL; t->Ffm = t->Ffp = NULL;
}

/* ------------------------------------------------------------------ writing */
static int put(FILE *f, const void *p, size_t n) { return fwrite(p, 1, n, f) != n; }

int kqed_write_taylor(const char *path, const double *rows, int nrow, int ncol)
{
  FILE *f = fopen(path, "wb");
  if (!f) { fprintf(stderr, "kqed: cannot write %s\n", path); return 1; }
  const int magic = KQED_MAGIC;
  uint32_t a = 0, b = 0;
  int e = put(f, &magic, 4) | put(f, &nrow, 4);
  for (int i = 0; i < nrow; i++) {
    e |= put(f, &ncol, 4) | put(f, rows + (size_t)i * ncol, 8u * ncol);
    kqed_cksum_accum(&a, &b, (uint32_t)i, rows + (size_t)i * ncol, 8u * ncol);
  }
  e |= put(f, &a, 4) | put(f, &b, 4);
  return (fclose(f) != 0) | e;
}

typedef void (*row_fn)(void *ctx, int nm, int ix, int iy, int nx, float *fm, float *fp);

static int write_ff_rows(const char *path, int nstpx, const double *XX, int nstpy, const double *YY,
                         const int *nfx, row_fn fn, void *ctx)
{
  FILE *f = fopen(path, "wb");
  if (!f) { fprintf(stderr, "kqed: cannot write %s\n", path); return 1; }
  setvbuf(f, NULL, _IOFBF, 1 << 22);
  const int magic = KQED_MAGIC, nffa = KQED_NFFA;
  uint32_t a, b;
  int e = put(f, &magic, 4) | put(f, &nstpx, 4) | put(f, XX, 8u * nstpx);
  a = b = 0; kqed_cksum_accum(&a, &b, 0, XX, 8u * nstpx);
  e |= put(f, &a, 4) | put(f, &b, 4);
  e |= put(f, &nstpy, 4) | put(f, YY, 8u * nstpy);
  a = b = 0; kqed_cksum_accum(&a, &b, 0, YY, 8u * nstpy);
  e |= put(f, &a, 4) | put(f, &b, 4) | put(f, &nffa, 4);
  int nmax = 0;
  for (int i = 0; i < nstpx; i++) if (nfx[i] > nmax) nmax = nfx[i];
  float *fm = (float *)malloc(4u * nmax), *fp = (float *)malloc(4u * nmax);
  uint32_t ma = 0, mb = 0, pa = 0, pb = 0, rank = 0;
  for (int nm = 0; nm < nffa; nm++) {
    e |= put(f, &nstpx, 4);
    for (int ix = 0; ix < nstpx; ix++) {
      e |= put(f, &nstpy, 4);
      for (int iy = 0; iy < nstpy; iy++) {
        const int nx = nfx[ix];
        fn(ctx, nm, ix, iy, nx, fm, fp);
        e |= put(f, &nx, 4) | put(f, fm, 4u * nx) | put(f, fp, 4u * nx);
        kqed_cksum_accum(&ma, &mb, rank, fm, 4u * nx);
        kqed_cksum_accum(&pa, &pb, rank, fp, 4u * nx);
        rank++;
      }
    }
  }
  e |= put(f, &ma, 4) | put(f, &mb, 4) | put(f, &pa, 4) | put(f, &pb, 4);
  free(fm); free(fp);
  e |= (fclose(f) != 0);
  if (e) fprintf(stderr, "kqed: write error on %s\n", path);
  return e;
}

static void copy_row(void *ctx, int nm, int ix, int iy, int nx, float *fm, float *fp)
{
  const kqed_host_tables *t = (const kqed_host_tables *)ctx;
  const size_t o = (((size_t)nm * t->nstpx + ix) * t->nstpy + iy) * t->nfmax;
  memcpy(fm, t->Ffm + o, 4u * nx);
  memcpy(fp, t->Ffp + o, 4u * nx);
}

int kqed_write_ff(const char *path, const kqed_host_tables *t)
{
  return write_ff_rows(path, t->nstpx, t->XX, t->nstpy, t->YY, t->nfx, copy_row, (void *)t);
}

/* ------------------------------------------------------------------ synthetic table */
static uint64_t mix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double u01(uint64_t seed, uint64_t a, uint64_t b)
{
  return (double)(mix64(seed ^ mix64(a * 1315423911ull + b)) >> 11) * (1.0 / 9007199254740992.0);
}

typedef struct {
  const kqed_synth_params *p;
  const double *XX, *YY;
  double amp[6], kap[6], lam[6];
  double *w;          /* [6][nfmax] */
  int nfmax;
} synth_ctx;

/* y-power exponent e_p of the reference's getff2 power case (getff-new.c:116-123):
 * yp = y^e_p, mshm; mm = |2 - mshm| */
static void power_case(int nm, int *ep, int *mm)
{
  const int base = nm < 6 ? nm : nm < 12 ? nm - 6 : (nm == 12 ? 2 : 3);
  switch (base) {
  case 2: *ep = -1; *mm = 1; break;  /* QG2: yp = 1/y, mshm = 3 */
  case 4: *ep = 1; *mm = 1; break;   /* QL4: yp = y,   mshm = 1 */
  case 5: *ep = -2; *mm = 2; break;  /* QL2: yp = 1/y^2, mshm = 4 */
  default: *ep = 0; *mm = 0; break;  /* QG0, QG1, QG3: yp = 1, mshm = 2 */
  }
}

static void synth_row(void *vctx, int nm, int ix, int iy, int nx, float *fm, float *fp)
{
  const synth_ctx *s = (const synth_ctx *)vctx;
  const int base = nm < 6 ? nm : nm < 12 ? nm - 6 : (nm == 12 ? 2 : 3);
  const int dord = nm < 6 ? 0 : nm < 12 ? 1 : 2;
  const double x = s->XX[ix], y = s->YY[iy];
  const int below = (x <= y);
  const double lrho = below ? log(x / y) : log(y / x);
  const double A = s->amp[base] * exp(-s->kap[base] * x - s->lam[base] * y);
  int ep, mm;
  power_case(nm, &ep, &mm);
  const double ly = log(y), lx = log(x);
  for (int j = 0; j < nx; j++) {
    double fac = 1.0;
    const double sj = below ? j / x : -j / x;
    const double q = below ? (double)j * (j - 1) / (x * x) : (double)j * (j + 1) / (x * x);
    const double k = s->kap[base];
    if (dord == 1) fac = -k + sj;
    else if (dord == 2) fac = k * k - 2.0 * k * sj + q;
    /* c_j = A w_j rho^j fac; fval = y^(ep-2-mm-j) Fm + y^(ep+mm+j) Fp */
    const double cj_over_rho = A * s->w[base * s->nfmax + j] * fac;
    double v;
    if (below) { /* Fm = c_j y^(2+mm+j-ep) = A w fac x^j y^(2+mm-ep) */
      v = cj_over_rho * exp(j * lx + (2 + mm - ep) * ly);
      fm[j] = (float)v; fp[j] = 0.0f;
    } else {     /* Fp = c_j y^-(ep+mm+j) = A w fac x^-j y^-(ep+mm) */
      v = cj_over_rho * exp(-j * lx - (ep + mm) * ly);
      fp[j] = (float)v; fm[j] = 0.0f;
    }
    (void)lrho;
  }
}

int kqed_synth_ff(const char *path, const kqed_synth_params *p)
{
  if (p->nstpx < 2 || p->nstpx > KQED_NTAYY_COLS || p->nstpy < 2 || p->nf0 < 4 || p->nf1 < p->nf0 ||
      p->nf1 > 4096 || !(p->xh > 0) || !(p->yh > 0)) {
    fprintf(stderr, "kqed: bad synthetic table parameters\n");
    return 1;
  }
  synth_ctx s;
  memset(&s, 0, sizeof s);
  s.p = p;
  double *XX = (double *)malloc(8u * p->nstpx), *YY = (double *)malloc(8u * p->nstpy);
  int *nfx = (int *)malloc(4u * p->nstpx);
  for (int i = 0; i < p->nstpx; i++) {
    /* x nodes on the Taylor-x y step: 3 steps apart up to node 14, 9 steps apart after it
     * (the layout the reference's Taylor-y table and regression outputs imply) */
    XX[i] = p->xh * (i <= 14 ? 3.0 * (i + 1) : 45.0 + 9.0 * (i - 14));
    nfx[i] = p->nf0 + (int)(((long)(p->nf1 - p->nf0) * i) / p->nstpx);
    /* Fp[j] grows like x^-j: keep it inside float range (the reason real tables carry fewer
     * Chebyshev terms at small x); the cap grows with x, so nfx stays non-decreasing */
    if (XX[i] < 1.0) {
      int cap = (int)(36.0 / log10(1.0 / XX[i]));
      if (cap < 4) cap = 4;
      if (nfx[i] > cap) nfx[i] = cap;
    }
  }
  for (int i = 0; i < p->nstpy; i++) YY[i] = p->yh * (i + 1);
  s.XX = XX; s.YY = YY;
  s.nfmax = p->nf1;
  s.w = (double *)malloc(sizeof(double) * 6 * s.nfmax);
  static const double amp[6] = { 2.0e-4, -1.5e-4, 1.0e-4, 3.0e-4, -2.0e-4, 1.2e-4 };
  for (int b = 0; b < 6; b++) {
    s.amp[b] = amp[b];
    s.kap[b] = 0.3 + 0.7 * u01(p->seed, 100 + b, 1);
    s.lam[b] = 0.2 + 0.6 * u01(p->seed, 200 + b, 2);
    const int alt = (b == 1 || b == 4);
    for (int j = 0; j < s.nfmax; j++) {
      const double u = u01(p->seed, 300 + b, (uint64_t)j);
      s.w[b * s.nfmax + j] = (0.5 + u) * pow(0.85, j) * ((alt && (j & 1)) ? -1.0 : 1.0);
    }
  }
  const int rc = write_ff_rows(path, p->nstpx, XX, p->nstpy, YY, nfx, synth_row, &s);
  free(XX); free(YY); free(nfx); free(s.w);
  return rc;
}

/* ------------------------------------------------------------------ ingestion checks (public) */

static void join(char *out, size_t n, const char *dir, const char *name) { snprintf(out, n, "%s/%s", dir, name); }

static int read_all(const char *dir, kqed_host_tables *t)
{
  char p[4096];
  memset(t, 0, sizeof *t);
  join(p, sizeof p, dir, "FFxy_single_cksum.bin");
  int rc = kqed_read_ff(p, t);
  if (!rc) { join(p, sizeof p, dir, "taylorx_cksum.bin"); rc = kqed_read_taylorx(p, t); }
  if (!rc) { join(p, sizeof p, dir, "taylory_cksum.bin"); rc = kqed_read_taylory(p, t); }
  return rc;
}

int kqed_tables_info(const char *dir, struct kqed_table_info *info)
{
  kqed_host_tables t;
  const int rc = read_all(dir, &t);
  if (!rc && info) {
    memset(info, 0, sizeof *info);
    info->nstpx = t.nstpx;
    info->nstpy = t.nstpy;
    info->nffa = KQED_NFFA;
    info->nfmax = t.nfmax;
    info->nfmin = t.nfmax;
    for (int i = 0; i < t.nstpx; i++) if (t.nfx[i] < info->nfmin) info->nfmin = t.nfx[i];
    info->x_first = t.XX[0];
    info->x_last = t.XX[t.nstpx - 1];
    info->y_first = t.YY[0];
    info->y_last = t.YY[t.nstpy - 1];
    info->x_increasing = info->y_increasing = 1;
    for (int i = 1; i < t.nstpx; i++) if (!(t.XX[i] > t.XX[i - 1])) info->x_increasing = 0;
    for (int i = 1; i < t.nstpy; i++) if (!(t.YY[i] > t.YY[i - 1])) info->y_increasing = 0;
    size_t bad = 0;
    const size_t nv = (size_t)KQED_NFFA * t.nstpx * t.nstpy * t.nfmax;
    for (size_t i = 0; i < nv; i++) bad += !isfinite(t.Ffm[i]) || !isfinite(t.Ffp[i]);
    info->nonfinite = bad;
  }
  kqed_free_host_tables(&t);
  return rc;
}

int kqed_tables_rewrite(const char *src_dir, const char *dst_dir)
{
  kqed_host_tables t;
  int rc = read_all(src_dir, &t);
  if (!rc) {
    char p[4096];
    join(p, sizeof p, dst_dir, "FFxy_single_cksum.bin");
    rc = kqed_write_ff(p, &t) ? 6 : 0;
    join(p, sizeof p, dst_dir, "taylorx_cksum.bin");
    if (!rc) rc = kqed_write_taylor(p, &t.TX[0][0], KQED_NTAYX_ROWS, KQED_NTAYX_COLS) ? 6 : 0;
    join(p, sizeof p, dst_dir, "taylory_cksum.bin");
    if (!rc) rc = kqed_write_taylor(p, &t.TY[0][0], KQED_NTAYY_ROWS, KQED_NTAYY_COLS) ? 6 : 0;
  }
  kqed_free_host_tables(&t);
  return rc;
}
