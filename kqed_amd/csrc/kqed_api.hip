/*
 * kqed_api.hip -- host side of the C ABI declared in include/kqed.h: table loading and device
 * residency, the batched pipeline driver, and the reference-compatible per-pair entry points.
 */
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <math.h>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/kqed.h"
#include "kqed_assembly.h"
#include "kqed_bank_layout.h"
#include "kqed_combine.h"
#include "kqed_device.h"
#include "kqed_pert.h"
#include "kqed_tables.h"

/* kernels (kqed_kernels.hip) */
__global__ void k_build_arrays(kqed_dev_tables T, const float *Fm, const float *Fp, double2 *P, double *PNC, double2 *S);
__global__ void k_fill_u32(unsigned *a, size_t n, unsigned v);
__global__ void k_classify(kqed_dev_tables T, size_t n, const double *x, int xs, const double *y, int ys, int mode,
                           unsigned *key, unsigned *rank, unsigned *gcount, int G, size_t nb);
__global__ void k_scan_tiles(unsigned *count, unsigned nb, unsigned *tsum, const unsigned *gcount, unsigned G, unsigned nkey);
__global__ void k_scan_add(unsigned *count, unsigned nb, const unsigned *tsum);
__global__ void k_scatter(size_t n, const unsigned *key, const unsigned *rank, const unsigned *start, unsigned *perm,
                          int G);
__global__ void k_algebra(size_t n, double *rec, const unsigned *kinds);
__global__ void k_inject_records(size_t n, double *rec, const unsigned *kinds, const double *w);
template <int CLS>
__global__ void k_interp(kqed_dev_tables T, size_t n, const double *x, int xs, const double *y, int ys,
                         const unsigned *perm, const unsigned *bounds, int ilo, int ihi, size_t c0, size_t c1, int which,
                         double M, double *rec, unsigned *kinds, const unsigned char *smask, int fuse);
__global__ void k_extract_scalars(kqed_dev_tables T, size_t n, const double *rec, const unsigned *kinds, double *out);
__global__ void k_surr_rowbin(kqed_dev_tables T, size_t n, const double *data, int *bin);
/* surrogate kernels (kqed_surrogate.hip) */
struct SurrRange { float lx0, lx1, ly0, ly1; };
struct SurrNet {
  const unsigned short *W[5];
  const float *b[5];
  const float *lsc;
  SurrRange R;
};
__global__ void k_gemm_bf16(int M, int N, int K, int kper, const float *A, long sam, long sak, const float *B, long sbk,
                            long sbn, float *C, long ldc, long zstride, const float *bias);
__global__ void k_sum_splits(long n, int nz, long zstride, const float *part, float *out);
__global__ void k_silu(long n, const float *Z, float *A);
__global__ void k_silu_bwd(long n, const float *Z, const float *dA, float *dZ);
__global__ void k_colsum_f(int B, int N, const float *dZ, float *db);
__global__ void k_mse(int B, int NP, int NOUT, const float *pred, const float *tgt, const float *wgt, float norm,
                      float *dpred, float *part);
__global__ void k_adam(long n, float *p, const float *g, float *m, float *v, float lr_t, float b1, float b2, float eps);
__global__ void k_surr_batch(int B, const unsigned *idx, const double *data, SurrRange R, const float *lsc, float *X,
                             float *Tg, float *W);
__global__ void k_surr_binmax(long n, const double *data, SurrRange R, unsigned *bmax, unsigned *bcount);
__global__ void k_surr_records(size_t n, double *rec, const unsigned *kinds, SurrNet net);
__global__ void k_surr_binerr(size_t n, const double *a, const double *b, const int *bin, unsigned *emax, unsigned *ecount,
                              float *err);
template <typename R>
__global__ void k_assemble(kqed_dev_tables T, size_t n, const double *rec, const unsigned *kinds, void *out, int layout,
                           int nslots, size_t ntot, const void *Tc, const double *Wc, const unsigned *split, size_t c0);
template <int AW, typename MT, bool WDOT>
__global__ void k_assemble2(kqed_dev_tables T, size_t n, const unsigned *split, size_t c0, const double *rec,
                            const unsigned *kinds, void *out, int layout, size_t ntot, const double *Wc);
template <int AW, typename MT>
__global__ void k_assemble2m(kqed_dev_tables T, size_t n, const unsigned *split, size_t c0, const double *rec,
                             const unsigned *kinds, void *out, int layout, int nslots, size_t ntot, const void *Tc);
template <int AW, typename MT>
__global__ void k_assemble2t(kqed_dev_tables T, size_t n, const unsigned *split, size_t c0, const double *rec,
                             const unsigned *kinds, void *out, int layout, int nslots, size_t ntot, const void *Tc);
__global__ void k_ragged_expand(size_t n, int ng, const unsigned long long *off, const double *y, double *ypair);
__global__ void k_contract_weights(size_t nx, const float *Tc, const unsigned *ent, const int *colx, double *W,
                                   double *colpart);
__global__ void k_block_sum(size_t n, size_t per, const double *in, double *out);
__global__ void k_colsum_final(int nb, const double *part, double *sum);
__global__ void k_combine(int ny, const double *sigma, const double *La, double ca, const double *sumT,
                          const double *consts, double cc0, double cc1, double *s);

#ifndef KQED_DEFAULT_TABLE_DIR
#define KQED_DEFAULT_TABLE_DIR "/usr/local/share/kqed/PRECOMP"
#endif

static thread_local char g_err[512] = "";
static void set_err(const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  fprintf(stderr, "kqed: %s\n", g_err);
}
extern "C" const char *kqed_last_error(void) { return g_err; }

#define HIPCHK(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) {                                                           \
      set_err("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
      goto fail;                                                                      \
    }                                                                                 \
  } while (0)

/* internal output modes: per-pair dot products with a site tensor (assembly kernel), and the
 * extracted interpolation results of the surrogate's training data (no algebra, no assembly) */
static const int LAYOUT_DOT = 16;
static const int LAYOUT_DOT64 = 64; /* per-pair dot with an FP64 tensor (passed as Tc) */
static const int LAYOUT_SCALARS = 32;

/* The surrogate as the pipeline sees it: network parameters for the inference kernel and the
 * per-bin mask of its error check (null = not validated: never used). */
struct SurrUse {
  SurrNet net;
  const unsigned char *mask;
  const double *inject = nullptr; /* weight functions from the caller instead of the network */
};

/* Device copies of the group offsets and the per-pair y vectors of a ragged batch. */
struct RaggedBufs {
  size_t n = 0, ng = 0;
  unsigned long long *off = nullptr;
  double *ypair = nullptr;
  void release()
  {
    (void)hipFree(off); (void)hipFree(ypair);
    off = nullptr; ypair = nullptr; n = ng = 0;
  }
};

/* sub-counters per sort bucket: a large batch counts each bucket in KQED_SORT_GROUPS counters
 * (the classification blocks take them in turn), which spreads the atomics on the buckets that most
 * blocks touch; the scan over (bucket, group) then gives every group its own start */
#define KQED_SORT_GROUPS 16
static int sort_groups(size_t n, int gmax)
{
  const int g = n >= ((size_t)1 << 22) ? 16 : n >= ((size_t)1 << 20) ? 4 : 1;
  return g < gmax ? g : gmax;
}

struct Workspace {
  size_t cap = 0, ncap = 0, nbuckets = 0;
  int groups = KQED_SORT_GROUPS; /* sub-counters per bucket that count holds */
  unsigned *key = nullptr, *rank = nullptr, *count = nullptr, *perm = nullptr, *kinds = nullptr;
  /* the classification's counters, groups-major (KQED_SORT_GROUPS arrays of nbuckets); null when
   * the sort runs with one group, which counts in count itself */
  unsigned *gcount = nullptr;
  double *rec = nullptr;
  /* contraction: per-pair dots [xcap], per-site pulled-back tensors W [xcap][KQED_WCOLS], and the small
   * reduction buffers (RED_LEN doubles + 4 * ycap * 384 + ...) */
  size_t xcap = 0, ycap = 0;
  double *dots = nullptr, *W = nullptr, *red = nullptr;
  /* kernel combinations: derived argument pairs [2][ccap][4] and their L0 tensors [ccap][384] */
  size_t ccap = 0;
  double *cargs = nullptr, *cL = nullptr;
  /* combination contraction: per-pair dots [NK][nx], y points, partial and final sums */
  size_t cdcap = 0;
  double *cdots = nullptr;
  /* lepton loop: iPihat [pcap][384], scaled pairs [2][pcap][4] */
  size_t pcap = 0;
  double *pP = nullptr, *pX = nullptr;
  /* ragged batches enqueued on this stream: group offsets and per-pair y vectors */
  RaggedBufs rb;
};

/* optional per-stage timing: events recorded on the launch stream around every kernel */
enum { ST_SORT = 0, ST_INTERP = 1, ST_ASSEMBLE = 2, ST_N = 3 };
struct StageEvent { int stage; hipEvent_t a, b; };

struct kqed_device {
  int device = 0;
  /* launch parameters: pairs per pipeline pass (bounds the record workspace), assembly grid,
   * sort key granularity and the batch size from which the sort is used */
  kqed_tune tune = { (size_t)1 << 21, 1024, KQED_SORT_CELL, 2048 };
  kqed_dev_tables dt;
  void *allocs[32];
  int nalloc = 0;
  kqed_map *maps = nullptr;
  /* the standard assembly map by column for the contraction pull-back: entry k of slot m of lane l
   * at csc_ent[64 (off_m + k) + l] as u32 (upper 16 bits of the FP64 coefficient << 16 | row), the
   * compact weight index of that column at csc_colx[64 m + l] (-1: no column) */
  unsigned *csc_ent = nullptr;
  int *csc_colx = nullptr;
  /* trace tables of the perturbative lepton loop (built on first use) */
  unsigned *pert_tab = nullptr;
  /* every surrogate bin set: routes all packed-path pairs through the substitution step */
  unsigned char *allmask = nullptr;
  /* device address of this GPU's copy of the allocation-free entry's scratch array (null when
   * hipGetSymbolAddress failed at set-up) */
  unsigned char *noalloc = nullptr;
  std::mutex mu;
  std::unordered_map<void *, Workspace> ws;
  bool profiling = false;
  std::vector<StageEvent> events;
  double stage_ms[ST_N] = { 0, 0, 0 };
  long stage_count[ST_N] = { 0, 0, 0 };
};

static void prof_begin(kqed_device *d, hipStream_t s, StageEvent &e, int stage)
{
  e.stage = stage;
  e.a = e.b = nullptr;
  if (!d->profiling) return;
  (void)hipEventCreate(&e.a);
  (void)hipEventCreate(&e.b);
  (void)hipEventRecord(e.a, s);
}
static void prof_end(kqed_device *d, hipStream_t s, StageEvent &e)
{
  if (!d->profiling || !e.a) return;
  (void)hipEventRecord(e.b, s);
  d->events.push_back(e);
}

extern "C" int kqed_profile(const struct QED_kernel_temps *t, int enable, double *ms, long *count)
{
  if (!t || !t->dev) return -1;
  kqed_device *d = t->dev;
  std::lock_guard<std::mutex> lk(d->mu);
  for (auto &e : d->events) {
    float f = 0.f;
    if (hipEventSynchronize(e.b) == hipSuccess && hipEventElapsedTime(&f, e.a, e.b) == hipSuccess) {
      d->stage_ms[e.stage] += f;
      d->stage_count[e.stage] += 1;
    }
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
  }
  d->events.clear();
  if (ms) for (int i = 0; i < ST_N; i++) ms[i] = d->stage_ms[i];
  if (count) for (int i = 0; i < ST_N; i++) count[i] = d->stage_count[i];
  for (int i = 0; i < ST_N; i++) { d->stage_ms[i] = 0.0; d->stage_count[i] = 0; }
  d->profiling = enable != 0;
  return 0;
}

static void *dev_alloc(kqed_device *d, size_t bytes)
{
  void *p = nullptr;
  if (d->nalloc >= (int)(sizeof d->allocs / sizeof d->allocs[0])) return nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  d->allocs[d->nalloc++] = p;
  return p;
}

static void free_ws(Workspace &w)
{
  (void)hipFree(w.key); (void)hipFree(w.rank); (void)hipFree(w.count); (void)hipFree(w.perm); (void)hipFree(w.kinds);
  (void)hipFree(w.gcount);
  (void)hipFree(w.rec); (void)hipFree(w.dots); (void)hipFree(w.W); (void)hipFree(w.red);
  (void)hipFree(w.cargs); (void)hipFree(w.cL); (void)hipFree(w.cdots);
  w.rb.release();
  (void)hipFree(w.pP); (void)hipFree(w.pX);
  w = Workspace();
}

extern "C" void free_QED_temps(struct QED_kernel_temps *t)
{
  if (!t) return;
  if (t->dev) {
    kqed_device *d = t->dev;
    (void)hipSetDevice(d->device);
    (void)hipDeviceSynchronize();
    for (auto &kv : d->ws) free_ws(kv.second);
    (void)hipFree(d->pert_tab);
    (void)hipFree(d->allmask);
    for (int i = 0; i < d->nalloc; i++) (void)hipFree(d->allocs[i]);
    free(d->maps);
    delete d;
  }
  free(t->Grid.XX);
  free(t->Grid.YY);
  free(t->Grid.nfx);
  free(t->G8);
  memset(t, 0, sizeof *t);
}

/* Deal the rows of each assembly map to the lanes of a wavefront (see kqed_device.h). */
static int build_wave_maps(kqed_device *d)
{
  kqed_dev_tables &T = d->dt;
  std::vector<unsigned> E;
  std::vector<unsigned short> R(3 * 384);
  /* the standard map's LDS layout (kqed_bank_layout.h): the slot of each tensor entry and the lane
   * of each row within its row slot, chosen to spread the gathers over the LDS banks */
  static const unsigned char std_slot[200] = KQED_STD_SLOT;
  static const unsigned short std_rows[384] = KQED_STD_ROWS;
  {
    bool seen_s[200] = {}, seen_r[384] = {};
    for (int c = 0; c < 200; c++) {
      if (std_slot[c] >= 200 || seen_s[std_slot[c]]) { set_err("bank layout: bad slot table"); return 1; }
      seen_s[std_slot[c]] = true;
    }
    for (int i = 0; i < 384; i++) {
      if (std_rows[i] >= 384 || seen_r[std_rows[i]]) { set_err("bank layout: bad row table"); return 1; }
      seen_r[std_rows[i]] = true;
    }
  }
  for (int q = 0; q < 3; q++) {
    const kqed_map &m = d->maps[q];
    int order[384];
    for (int r = 0; r < 384; r++) order[r] = r;
    std::stable_sort(order, order + 384, [&](int a, int b) {
      return m.rowptr[a + 1] - m.rowptr[a] > m.rowptr[b + 1] - m.rowptr[b];
    });
    if (q == 0) {
      /* same rows per row slot (a permutation within each slot), in the layout's lane order */
      for (int sl = 0; sl < 6; sl++) {
        std::vector<int> a(order + 64 * sl, order + 64 * sl + 64), b(std_rows + 64 * sl, std_rows + 64 * sl + 64);
        std::sort(a.begin(), a.end());
        std::sort(b.begin(), b.end());
        if (a != b) { set_err("bank layout does not match the standard map"); return 1; }
      }
      for (int i = 0; i < 384; i++) order[i] = std_rows[i];
    }
    static const int KB[6] = KQED_WM_K;
    for (int sl = 0; sl < 6; sl++) {
      T.wmOff[q][sl] = (int)E.size();
      int K = 0;
      for (int l = 0; l < 64; l++) {
        const int r = order[64 * sl + l];
        K = std::max(K, m.rowptr[r + 1] - m.rowptr[r]);
        R[q * 384 + 64 * sl + l] = (unsigned short)r;
      }
      /* the kernel unrolls each slot to its compile-time length; shorter rows take zero entries */
      if (K > KB[sl]) { set_err("assembly map row longer than the kernel's unrolled length"); return 1; }
      K = KB[sl];
      T.wmK[q][sl] = K;
      for (int k = 0; k < K; k++)
        for (int l = 0; l < 64; l++) {
          const int r = order[64 * sl + l], b = m.rowptr[r], len = m.rowptr[r + 1] - b;
          /* entry: byte offset of the column's slot in a row of two-pair FP64 slots (slot * 16, the
           * address the two-pair kernels use as it is) in the low half, the
           * upper 16 bits of the coefficient's FP64 bit pattern (+-0.5, +-1, +-2: low 48 bits are
           * zero) in the high half; padding takes the zero slot (col 200) */
          double v = 1.0;
          int col = 200;
          if (k < len) {
            v = m.val[b + k];
            col = m.col[b + k];
            const double a = fabs(v);
            if ((a != 0.5 && a != 1.0 && a != 2.0) || col >= 200) { set_err("assembly coefficient %g not encodable", v); return 1; }
          }
          const int slot = q == 0 && col < 200 ? std_slot[col] : col;
          unsigned long long bits;
          memcpy(&bits, &v, 8);
          E.push_back((unsigned)(bits >> 48) << 16 | (unsigned)(slot * 16));
        }
    }
    T.wmOff[q][6] = (int)E.size();
    for (int sl = 0, cum = 0; sl <= 6; cum += sl < 6 ? KB[sl] : 0, sl++)
      if (T.wmOff[q][sl] != 64 * (KQED_WM_PER_LANE * q + cum)) { set_err("wave map layout mismatch"); return 1; }
  }
  T.wmTotal = (int)E.size();
  if (T.wmTotal > KQED_WM_MAX) { set_err("assembly maps too large for LDS"); return 1; }
  unsigned *e = (unsigned *)dev_alloc(d, 4 * E.size());
  unsigned short *r = (unsigned short *)dev_alloc(d, 2 * R.size());
  unsigned char *sl = (unsigned char *)dev_alloc(d, 256);
  if (!e || !r || !sl || hipMemcpy(e, E.data(), 4 * E.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(r, R.data(), 2 * R.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(sl, std_slot, 200, hipMemcpyHostToDevice) != hipSuccess) {
    set_err("wave map upload failed");
    return 1;
  }
  T.wmE = e;
  T.wmRow = r;
  T.wmSlot = sl;
  return 0;
}

/* ---------------------------------------------------------------- allocation-free scratch */
/* Scratch of kqed_batch_noalloc, one copy per GPU for the life of the process: the records and
 * kinds of one pass of KQED_NOALLOC_PASS pairs, the sort keys, ranks and permutation of
 * KQED_NOALLOC_MAX_PAIRS pairs and KQED_NOALLOC_MAX_BUCKETS sort-bucket counters (every part
 * 256-byte aligned).  Host code never uses the name `scratch` as an address: initialise() takes
 * this GPU's device address with hipGetSymbolAddress and keeps it in kqed_device::noalloc. */
static constexpr size_t al256(size_t b) { return (b + 255) / 256 * 256; }
static constexpr size_t NA_REC = 0;
static constexpr size_t NA_KINDS = NA_REC + al256((size_t)KQED_NOALLOC_PASS * KQED_REC_LEN * 8);
static constexpr size_t NA_KEY = NA_KINDS + al256((size_t)KQED_NOALLOC_PASS * 4);
static constexpr size_t NA_RANK = NA_KEY + al256((size_t)KQED_NOALLOC_MAX_PAIRS * 4);
static constexpr size_t NA_PERM = NA_RANK + al256((size_t)KQED_NOALLOC_MAX_PAIRS * 4);
static constexpr size_t NA_COUNT = NA_PERM + al256((size_t)KQED_NOALLOC_MAX_PAIRS * 4);
static constexpr size_t NA_BYTES = NA_COUNT + al256(((size_t)KQED_NOALLOC_MAX_BUCKETS + (KQED_NOALLOC_MAX_BUCKETS + 1023) / 1024) * 4);
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

extern "C" int kqed_initialise_dir(struct QED_kernel_temps *t, const char *dir)
{
  memset(t, 0, sizeof *t);
  kqed_host_tables *h = (kqed_host_tables *)calloc(1, sizeof(kqed_host_tables));
  kqed_device *d = new kqed_device();
  char path[4096];
  float *dFm = nullptr, *dFp = nullptr;
  int ndev = 0;
  if (!h) { set_err("out of host memory"); delete d; return 1; }
  snprintf(path, sizeof path, "%s/FFxy_single_cksum.bin", dir);
  if (kqed_read_ff(path, h)) { set_err("cannot load %s", path); goto fail_host; }
  snprintf(path, sizeof path, "%s/taylorx_cksum.bin", dir);
  if (kqed_read_taylorx(path, h)) { set_err("cannot load %s", path); goto fail_host; }
  snprintf(path, sizeof path, "%s/taylory_cksum.bin", dir);
  if (kqed_read_taylory(path, h)) { set_err("cannot load %s", path); goto fail_host; }
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    set_err("no HIP device available (this library runs its kernels on the GPU only)");
    goto fail_host;
  }
  HIPCHK(hipGetDevice(&d->device));
  /* this device's copy of the allocation-free entry's scratch (the device is current here) */
  if (hipGetSymbolAddress((void **)&d->noalloc, HIP_SYMBOL(scratch)) != hipSuccess) {
    d->noalloc = nullptr;
    (void)hipGetLastError();
  }
  {
    kqed_dev_tables &T = d->dt;
    T.nstpx = h->nstpx;
    T.nstpy = h->nstpy;
    T.nfmax = h->nfmax;
    T.ks = h->nfmax + 2 + KQED_CLEN_BLOCK - 1;
    /* the packed-path kernel stages node arrays in LDS: KQED_LDS_ARRAYS of ks double2 per wavefront */
    if ((size_t)KQED_INTERP_THREADS / 64 * KQED_LDS_ARRAYS * T.ks * sizeof(double2) > 64 * 1024) {
      set_err("table nfmax %d too large for the packed-path LDS staging (at most %d)", h->nfmax,
              (int)(64 * 1024 / (KQED_INTERP_THREADS / 64 * KQED_LDS_ARRAYS * sizeof(double2))) - 2 - KQED_CLEN_BLOCK + 1);
      goto fail;
    }
    const size_t nodes = (size_t)h->nstpx * h->nstpy;
    const size_t nraw = (size_t)KQED_NFFA * nodes * h->nfmax;
    double *XX = (double *)dev_alloc(d, 8 * h->nstpx), *YY = (double *)dev_alloc(d, 8 * h->nstpy);
    int *nfx = (int *)dev_alloc(d, 4 * h->nstpx);
    double *TX = (double *)dev_alloc(d, sizeof h->TX), *TY = (double *)dev_alloc(d, sizeof h->TY);
    double *LUT = (double *)dev_alloc(d, 8 * 8 * (size_t)T.ks);
    double2 *P = (double2 *)dev_alloc(d, 16 * KQED_NFFA * nodes * T.ks);
    double *PNC = (double *)dev_alloc(d, 8 * 4 * nodes * T.ks);
    double2 *S = (double2 *)dev_alloc(d, 16 * KQED_NFFA * nodes * T.ks);
    if (!XX || !YY || !nfx || !TX || !TY || !LUT || !P || !PNC || !S) { set_err("device allocation failed"); goto fail; }
    HIPCHK(hipMemcpy(XX, h->XX, 8 * h->nstpx, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(YY, h->YY, 8 * h->nstpy, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(nfx, h->nfx, 4 * h->nstpx, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(TX, h->TX, sizeof h->TX, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(TY, h->TY, sizeof h->TY, hipMemcpyHostToDevice));
    /* ratio LUTs of the derivative recurrences (reference simd.c:716-784) */
    double *lut = (double *)calloc(8 * (size_t)T.ks, sizeof(double));
    for (int i = 1; i < T.ks; i++) { lut[8 * i + 1] = (i + 1) / (double)i; lut[8 * i + 2] = (i + 3) / (double)(i + 1); }
    for (int i = 2; i < T.ks; i++) { lut[8 * i + 3] = (i + 1) / (double)(i - 1); lut[8 * i + 4] = (i + 4) / (double)i; }
    for (int i = 3; i < T.ks; i++) { lut[8 * i + 5] = (i + 1) / (double)(i - 2); lut[8 * i + 6] = (i + 5) / (double)(i - 1); }
    hipError_t e = hipMemcpy(LUT, lut, 8 * 8 * (size_t)T.ks, hipMemcpyHostToDevice);
    free(lut);
    HIPCHK(e);
    T.XX = XX; T.YY = YY; T.nfx = nfx; T.TX = TX; T.TY = TY; T.LUT = LUT; T.P = P; T.PNC = PNC; T.S = S;
    /* raw single-precision tables only live until the per-node arrays are built */
    HIPCHK(hipMalloc(&dFm, 4 * nraw));
    HIPCHK(hipMalloc(&dFp, 4 * nraw));
    HIPCHK(hipMemcpy(dFm, h->Ffm, 4 * nraw, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dFp, h->Ffp, 4 * nraw, hipMemcpyHostToDevice));
    const size_t nthr = KQED_NFFA * nodes;
    hipLaunchKernelGGL(k_build_arrays, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, 0, T, dFm, dFp, P, PNC, S);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    (void)hipFree(dFm); dFm = nullptr;
    (void)hipFree(dFp); dFp = nullptr;
    /* assembly maps from the Dirac traces */
    d->maps = (kqed_map *)calloc(3, sizeof(kqed_map));
    if (!d->maps || kqed_build_maps(&d->maps[0], &d->maps[1], &d->maps[2])) { set_err("assembly map build failed"); goto fail; }
    if (d->maps[0].kmax > 14 || d->maps[1].kmax > 10 || d->maps[2].kmax > 10) { set_err("assembly map rows too long"); goto fail; }
    if (build_wave_maps(d)) goto fail;
    {
      /* standard map by column: entries of column c in row order */
      const kqed_map &m0 = d->maps[0];
      std::vector<int> cp(201, 0);
      std::vector<unsigned> ent(m0.nnz);
      for (int e = 0; e < m0.nnz; e++) cp[m0.col[e] + 1]++;
      for (int c = 0; c < 200; c++) cp[c + 1] += cp[c];
      std::vector<int> fill(cp.begin(), cp.end() - 1);
      for (int r = 0; r < 384; r++)
        for (int e = m0.rowptr[r]; e < m0.rowptr[r + 1]; e++) {
          const int k = fill[m0.col[e]]++;
          unsigned long long bits;
          memcpy(&bits, &m0.val[e], 8);
          if (bits & 0xFFFFFFFFFFFFull) { set_err("standard map coefficient %g not encodable", m0.val[e]); goto fail; }
          ent[k] = (unsigned)(bits >> 48) << 16 | (unsigned)r;
        }
      /* columns (those a row reads: exactly the ones with a compact weight index) sorted by length,
       * longest first, and dealt to lane l of slot m in the order 64 m + l; each slot is unrolled to
       * the length of its longest column (padding entries have coefficient 0) */
      std::vector<int> order;
      for (int c = 0; c < 200; c++) {
        if ((cp[c + 1] > cp[c]) != (kqed_wcol(c) >= 0)) { set_err("standard map columns differ from the weight layout"); goto fail; }
        if (cp[c + 1] > cp[c]) order.push_back(c);
      }
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cp[a + 1] - cp[a] > cp[b + 1] - cp[b]; });
      {
        const int K[3] = KQED_PB_K;
        std::vector<unsigned> ent2(64 * KQED_PB_TOTAL, 0u);
        std::vector<int> colx(3 * 64, -1);
        int off = 0;
        for (int m = 0; m < 3; m++) {
          const int len0 = 64 * m < (int)order.size() ? cp[order[64 * m] + 1] - cp[order[64 * m]] : 0;
          if (len0 != K[m] || (m == 2 && (int)order.size() > 192)) { set_err("pull-back slot lengths differ from the kernel's"); goto fail; }
          for (int l = 0; l < 64 && 64 * m + l < (int)order.size(); l++) {
            const int c = order[64 * m + l];
            colx[64 * m + l] = kqed_wcol(c);
            for (int k = 0; k < cp[c + 1] - cp[c]; k++) ent2[64 * (off + k) + l] = ent[cp[c] + k];
          }
          off += K[m];
        }
        d->csc_ent = (unsigned *)dev_alloc(d, 4 * ent2.size());
        d->csc_colx = (int *)dev_alloc(d, 4 * colx.size());
        if (!d->csc_ent || !d->csc_colx) { set_err("device allocation failed"); goto fail; }
        HIPCHK(hipMemcpy(d->csc_ent, ent2.data(), 4 * ent2.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d->csc_colx, colx.data(), 4 * colx.size(), hipMemcpyHostToDevice));
      }
    }
  }
  /* host-visible mirror of the grid (reference Grid_coeffs fields) */
  t->Grid.Nffa = KQED_NFFA;
  t->Grid.nstpx = h->nstpx;
  t->Grid.nstpy = h->nstpy;
  t->Grid.XX = h->XX; h->XX = nullptr;
  t->Grid.YY = h->YY; h->YY = nullptr;
  t->Grid.nfx = h->nfx; h->nfx = nullptr;
  t->Grid.xstp = t->Grid.XX[1] - t->Grid.XX[0];
  t->Grid.ystp = t->Grid.YY[1] - t->Grid.YY[0];
  t->Grid.NtayY = KQED_NTAYX_ROWS; t->Grid.NY_tay = KQED_NTAYX_COLS;
  t->Grid.NtayX = KQED_NTAYY_ROWS; t->Grid.NX_tay = KQED_NTAYY_COLS;
  t->G8 = (int *)malloc(65536 * sizeof(int));
  if (t->G8) kqed_fill_g8(t->G8);
  t->dev = d;
  kqed_free_host_tables(h);
  free(h);
  return 0;
fail:
  if (dFm) (void)hipFree(dFm);
  if (dFp) (void)hipFree(dFp);
  for (int i = 0; i < d->nalloc; i++) (void)hipFree(d->allocs[i]);
  free(d->maps);
fail_host:
  delete d;
  kqed_free_host_tables(h);
  free(h);
  memset(t, 0, sizeof *t);
  return 1;
}

extern "C" int initialise(struct QED_kernel_temps *t)
{
  const char *dir = getenv("KQED_TABLE_DIR");
  return kqed_initialise_dir(t, dir && *dir ? dir : KQED_DEFAULT_TABLE_DIR);
}

/* pass records and kinds of min(n, chunk) pairs, sort keys/ranks/permutation of n pairs and one
 * counter per sort bucket (special pairs, nstpx * nstpy cells, nstpy Taylor-region y nodes) */
static size_t batch_workspace(size_t n, size_t chunk, size_t nbuckets)
{
  const size_t c = n < chunk ? n : chunk;
  const size_t nc = nbuckets * KQED_SORT_GROUPS;
  return c * (KQED_REC_LEN * 8 + 4) + 12 * n + 4 * (2 * nc + (nc + 1023) / 1024);
}

extern "C" size_t kqed_batch_workspace(size_t n)
{
  /* the default pass of 2^21 pairs and the largest table the reference's formats allow here
   * (100 x nodes, 4096 y nodes) */
  return batch_workspace(n, (size_t)1 << 21, 1 + 100 * 4096 + 4096);
}

extern "C" size_t kqed_batch_workspace_tables(const struct QED_kernel_temps *t, size_t n)
{
  if (!t || !t->dev) return 0;
  kqed_device *d = t->dev;
  std::lock_guard<std::mutex> lk(d->mu);
  return batch_workspace(n, d->tune.chunk, 1 + (size_t)d->dt.nstpx * d->dt.nstpy + d->dt.nstpy);
}

/* records for one pass of c pairs; sort keys and permutation for the whole batch of n pairs */
static int ensure_ws(kqed_device *d, Workspace &w, size_t c, size_t n)
{
  const size_t nb = 1 + (size_t)d->dt.nstpx * d->dt.nstpy + d->dt.nstpy;
  if (w.cap < c || w.nbuckets != nb) {
    (void)hipFree(w.kinds); (void)hipFree(w.rec); (void)hipFree(w.count); (void)hipFree(w.gcount);
    w.kinds = w.count = w.gcount = nullptr; w.rec = nullptr; w.cap = 0;
    const size_t nc = nb * KQED_SORT_GROUPS;
    if (hipMalloc(&w.kinds, 4 * c) != hipSuccess || hipMalloc(&w.count, 4 * (nc + (nc + 1023) / 1024)) != hipSuccess ||
        hipMalloc(&w.gcount, 4 * nc) != hipSuccess || hipMalloc(&w.rec, 8 * KQED_REC_LEN * c) != hipSuccess) {
      set_err("workspace allocation of %zu pairs failed", c);
      return -2;
    }
    w.cap = c;
    w.nbuckets = nb;
  }
  if (w.ncap < n) {
    (void)hipFree(w.key); (void)hipFree(w.rank); (void)hipFree(w.perm);
    w.key = w.rank = w.perm = nullptr; w.ncap = 0;
    if (hipMalloc(&w.key, 4 * n) != hipSuccess || hipMalloc(&w.rank, 4 * n) != hipSuccess ||
        hipMalloc(&w.perm, 4 * n) != hipSuccess) {
      set_err("sort workspace allocation of %zu pairs failed", n);
      return -2;
    }
    w.ncap = n;
  }
  return 0;
}

/* The pipeline over n pairs (x[xs*i..], y[ys*i..]), lock held: one sort of the whole batch by
 * regime and interpolation cell, then passes of tune.chunk pairs (interpolation per regime class,
 * chain-rule algebra, assembly).  Only kernel launches: it can be captured into a graph once the
 * workspace is sized (ensure_ws with the same sizes beforehand). */
static int run_pairs(kqed_device *d, Workspace &w, hipStream_t s, int which, double M, size_t n, const double *x_dev,
                     int xs, const double *y_dev, int ys, void *out_dev, int layout, const void *Tc,
                     const double *Wc, const SurrUse *sur = nullptr)
{
  if (n == 0) return 0;
  const kqed_tune &tu = d->tune;
  const size_t CH = tu.chunk;
  const bool sort = n >= tu.sort_min && tu.sort_mode != KQED_SORT_NONE;
  if (ensure_ws(d, w, n < CH ? n : CH, sort ? n : 0)) return -2;
  const kqed_dev_tables &T = d->dt;
  const unsigned *perm = nullptr;
  int G = 1; /* sub-counters per bucket: the start of bucket b is count[b * G] */
  StageEvent ev;
  if (sort) {
    /* one counting sort of the whole batch by interpolation cell; passes then walk the sorted
     * order, so every pass sees the densest possible cells */
    const unsigned blocks = (unsigned)((n + 255) / 256);
    prof_begin(d, s, ev, ST_SORT);
    G = sort_groups(n, w.gcount ? w.groups : 1);
    const size_t nc = w.nbuckets * G;
    unsigned *gc = G > 1 ? w.gcount : w.count;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, s, gc, nc, 0u);
    hipLaunchKernelGGL(k_classify, dim3(blocks), dim3(256), 0, s, T, n, x_dev, xs, y_dev, ys, tu.sort_mode, w.key, w.rank,
                       gc, G, w.nbuckets);
    const unsigned ntile = (unsigned)((nc + 1023) / 1024);
    hipLaunchKernelGGL(k_scan_tiles, dim3(ntile), dim3(1024), 0, s, w.count, (unsigned)nc, w.count + nc, (const unsigned *)gc,
                       (unsigned)G, (unsigned)w.nbuckets);
    hipLaunchKernelGGL(k_scan_add, dim3(ntile), dim3(1024), 0, s, w.count, (unsigned)nc, w.count + nc);
    hipLaunchKernelGGL(k_scatter, dim3(blocks), dim3(256), 0, s, n, w.key, w.rank, w.count, w.perm, G);
    perm = w.perm;
    prof_end(d, s, ev);
  }
  static const int nslots[5] = { 1, 3, 3, 4, 3 };
  /* the interpolation kernels run the chain-rule algebra themselves unless the stage-1 values are
   * the output (extraction); records whose values a surrogate or the caller supplies (flagged
   * KQED_SURR_BIT) get theirs in k_algebra after the substitution */
  const int fuse = layout != LAYOUT_SCALARS;
  for (size_t c0 = 0; c0 < n; c0 += CH) {
    const size_t c = n - c0 < CH ? n - c0 : CH, c1 = c0 + c;
    const unsigned blocks = (unsigned)((c + 255) / 256);
    const unsigned iblocks = (unsigned)((c + KQED_INTERP_THREADS - 1) / KQED_INTERP_THREADS);
    prof_begin(d, s, ev, ST_INTERP);
    if (perm) {
      /* count[b G] is the start of bucket b: special pairs [0, count[G]), packed path
       * [count[G], count[(1 + nstpx*nstpy) G]), Taylor-x region [count[(1 + nstpx*nstpy) G], n) */
      const int ip = (1 + T.nstpx * T.nstpy) * G;
      const unsigned char *smask = sur ? sur->mask : nullptr;
      hipLaunchKernelGGL(k_interp<0>, dim3(iblocks), dim3(KQED_INTERP_THREADS), 0, s, T, n, x_dev, xs, y_dev, ys, perm, w.count, -1, G, c0, c1,
                         which, M, w.rec, w.kinds, (const unsigned char *)nullptr, fuse);
      hipLaunchKernelGGL(k_interp<1>, dim3(iblocks), dim3(KQED_INTERP_THREADS), KQED_INTERP_THREADS / 64 * KQED_LDS_ARRAYS * T.ks * sizeof(double2), s, T, n, x_dev, xs, y_dev, ys, perm, w.count, G, ip, c0, c1,
                         which, M, w.rec, w.kinds, smask, fuse);
      hipLaunchKernelGGL(k_interp<2>, dim3(iblocks), dim3(KQED_INTERP_THREADS), 0, s, T, n, x_dev, xs, y_dev, ys, perm, w.count, ip, -1, c0, c1,
                         which, M, w.rec, w.kinds, (const unsigned char *)nullptr, fuse);
      if (smask && sur->inject)
        hipLaunchKernelGGL(k_inject_records, dim3(blocks), dim3(256), 0, s, c, w.rec, w.kinds, sur->inject);
      else if (smask)
        hipLaunchKernelGGL(k_surr_records, dim3((unsigned)((c + 63) / 64)), dim3(256), 0, s, c, w.rec, w.kinds, sur->net);
    } else {
      hipLaunchKernelGGL(k_interp<0>, dim3(iblocks), dim3(KQED_INTERP_THREADS), 0, s, T, n, x_dev, xs, y_dev, ys, perm, w.count, -1, -1, c0, c1,
                         which, M, w.rec, w.kinds, (const unsigned char *)nullptr, fuse);
    }
    if (layout == LAYOUT_SCALARS) {
      hipLaunchKernelGGL(k_extract_scalars, dim3(blocks), dim3(256), 0, s, T, c, w.rec, w.kinds, (double *)out_dev);
      prof_end(d, s, ev);
      continue;
    }
    if (sur) hipLaunchKernelGGL(k_algebra, dim3(blocks), dim3(256), 0, s, c, w.rec, w.kinds);
    prof_end(d, s, ev);
    const size_t nblk = (c + KQED_ASM_WAVES - 1) / KQED_ASM_WAVES;
    const unsigned grid = (unsigned)(nblk < (size_t)tu.asm_grid ? nblk : (size_t)tu.asm_grid);
    prof_begin(d, s, ev, ST_ASSEMBLE);
    /* one evaluation slot in sorted order: the records past the special pairs (kinds KIND_STD /
     * KIND_NONE only) go to the two-pair kernel (outputs, or the contraction shortcut's dot), the
     * special ones to the general one */
    const bool f32 = layout == KQED_AOS_F32 || layout == KQED_SOA_F32;
    const bool two = perm && nslots[which] == 1 &&
                     (((f32 || layout == KQED_AOS || layout == KQED_SOA) && !Wc) || (Wc && layout == LAYOUT_DOT));
    /* several slots (L1..L3, M-L2): the multi-slot two-pair kernel, which also forms the
     * contraction dot */
    /* a one-slot dot pass without the linear shortcut (the contraction's x-only sums, whose pairs are
     * all special points): every record, special or not, two per wavefront in the multi-slot kernel */
    const bool dot_all = perm && nslots[which] == 1 && !Wc && layout == LAYOUT_DOT;
    const bool twom = perm && !Wc &&
                      ((nslots[which] > 1 && (f32 || layout == KQED_AOS || layout == KQED_SOA || layout == LAYOUT_DOT ||
                                              layout == LAYOUT_DOT64)) || dot_all);
    /* split[0]: the first record the two-pair kernels take (count[G]: past the special pairs; count[0]
     * = 0: every record); the one-pair kernel takes the records before it */
    const unsigned *split = dot_all ? w.count : two || twom ? w.count + G : nullptr;
    if (f32)
      hipLaunchKernelGGL(k_assemble<float>, dim3(grid), dim3(64 * KQED_ASM_WAVES), 0, s, T, c, w.rec, w.kinds, out_dev, layout,
                         nslots[which], n, Tc, Wc, split, c0);
    else
      hipLaunchKernelGGL(k_assemble<double>, dim3(grid), dim3(64 * KQED_ASM_WAVES), 0, s, T, c, w.rec, w.kinds, out_dev, layout,
                         nslots[which], n, Tc, Wc, split, c0);
    if (two) {
      const size_t nb2 = ((c + 1) / 2 + KQED_ASM2_WAVES - 1) / KQED_ASM2_WAVES;
      const unsigned grid2 = (unsigned)(nb2 < (size_t)tu.asm_grid ? nb2 : (size_t)tu.asm_grid);
      const size_t nb2d = ((c + 1) / 2 + KQED_ASM2D_WAVES - 1) / KQED_ASM2D_WAVES;
      const unsigned grid2d = (unsigned)(nb2d < (size_t)tu.asm_grid ? nb2d : (size_t)tu.asm_grid);
      if (f32)
        hipLaunchKernelGGL((k_assemble2<KQED_ASM2_WAVES, float, false>), dim3(grid2), dim3(64 * KQED_ASM2_WAVES), 0, s, T, c,
                           split, c0, w.rec, w.kinds, out_dev, layout, n, (const double *)nullptr);
      else if (Wc)
        hipLaunchKernelGGL((k_assemble2<KQED_ASM2D_WAVES, double, true>), dim3(grid2d), dim3(64 * KQED_ASM2D_WAVES), 0, s, T, c,
                           split, c0, w.rec, w.kinds, out_dev, layout, n, Wc);
      else
        hipLaunchKernelGGL((k_assemble2<KQED_ASM2_WAVES, double, false>), dim3(grid2), dim3(64 * KQED_ASM2_WAVES), 0, s, T, c,
                           split, c0, w.rec, w.kinds, out_dev, layout, n, (const double *)nullptr);
    }
    if (twom && ((which >= 1 && which <= 3) || dot_all)) {
      /* L1, L2, L3 and the one-slot dot passes over special points: one Tabd map application on the
       * combined inputs */
      const size_t nb2 = ((c + 1) / 2 + KQED_ASM2T_WAVES - 1) / KQED_ASM2T_WAVES;
      const unsigned grid2 = (unsigned)(nb2 < (size_t)tu.asm_grid ? nb2 : (size_t)tu.asm_grid);
      if (f32)
        hipLaunchKernelGGL((k_assemble2t<KQED_ASM2T_WAVES, float>), dim3(grid2), dim3(64 * KQED_ASM2T_WAVES), 0, s, T, c,
                           split, c0, w.rec, w.kinds, out_dev, layout, nslots[which], n, Tc);
      else
        hipLaunchKernelGGL((k_assemble2t<KQED_ASM2T_WAVES, double>), dim3(grid2), dim3(64 * KQED_ASM2T_WAVES), 0, s, T, c,
                           split, c0, w.rec, w.kinds, out_dev, layout, nslots[which], n, Tc);
    } else if (twom) {
      const size_t nb2 = ((c + 1) / 2 + KQED_ASM2M_WAVES - 1) / KQED_ASM2M_WAVES;
      const unsigned grid2 = (unsigned)(nb2 < (size_t)tu.asm_grid ? nb2 : (size_t)tu.asm_grid);
      if (f32)
        hipLaunchKernelGGL((k_assemble2m<KQED_ASM2M_WAVES, float>), dim3(grid2), dim3(64 * KQED_ASM2M_WAVES), 0, s, T, c,
                           split, c0, w.rec, w.kinds, out_dev, layout, nslots[which], n, Tc);
      else
        hipLaunchKernelGGL((k_assemble2m<KQED_ASM2M_WAVES, double>), dim3(grid2), dim3(64 * KQED_ASM2M_WAVES), 0, s, T, c,
                           split, c0, w.rec, w.kinds, out_dev, layout, nslots[which], n, Tc);
    }
    prof_end(d, s, ev);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) { set_err("kernel launch failed: %s", hipGetErrorString(e)); return -3; }
  return 0;
}

static bool layout_ok(int layout)
{
  return layout == KQED_AOS || layout == KQED_SOA || layout == KQED_AOS_F32 || layout == KQED_SOA_F32;
}

extern "C" int kqed_batch(const struct QED_kernel_temps *t, int which, double M, size_t n, const double *x_dev,
                          const double *y_dev, double *out_dev, int layout, void *stream)
{
  if (!t || !t->dev) { set_err("kqed_batch: tables not initialised"); return -1; }
  if (which < 0 || which > 4 || !layout_ok(layout)) { set_err("kqed_batch: bad argument"); return -1; }
  if (n == 0) return 0;
  if (n > 0xFFFFFFFFull) { set_err("kqed_batch: at most 2^32-1 pairs per call"); return -1; }
  kqed_device *d = t->dev;
  std::lock_guard<std::mutex> lk(d->mu);
  return run_pairs(d, d->ws[stream], (hipStream_t)stream, which, M, n, x_dev, 4, y_dev, 4, out_dev, layout, nullptr,
                   nullptr);
}

/* kqed_batch without any device allocation: the pass workspace is this GPU's scratch array
 * (kqed_device::noalloc), passes of at most KQED_NOALLOC_PASS pairs.  Outputs equal kqed_batch's
 * (results do not depend on the pass length).  Calls that overlap on one GPU share the array and
 * race; the pipeline clears its counters itself (fill kernel) on every call. */
extern "C" int kqed_batch_noalloc(const struct QED_kernel_temps *t, int which, double M, size_t n, const double *x_dev,
                                  const double *y_dev, void *out_dev, int layout, void *stream)
{
  if (!t || !t->dev) { set_err("kqed_batch_noalloc: tables not initialised"); return -1; }
  if (which < 0 || which > 4 || !layout_ok(layout)) { set_err("kqed_batch_noalloc: bad argument"); return -1; }
  if (n == 0) return 0;
  if (n > (size_t)KQED_NOALLOC_MAX_PAIRS) {
    set_err("kqed_batch_noalloc: %zu pairs exceed the maximum of %d per call", n, KQED_NOALLOC_MAX_PAIRS);
    return -1;
  }
  kqed_device *d = t->dev;
  if (!d->noalloc) { set_err("kqed_batch_noalloc: scratch address unavailable on this device"); return -3; }
  const size_t nb = 1 + (size_t)d->dt.nstpx * d->dt.nstpy + d->dt.nstpy;
  if (nb > (size_t)KQED_NOALLOC_MAX_BUCKETS) {
    set_err("kqed_batch_noalloc: the table's %zu sort buckets exceed the maximum of %d", nb, KQED_NOALLOC_MAX_BUCKETS);
    return -1;
  }
  Workspace w; /* a view of the scratch array, sized so that the pipeline allocates nothing */
  w.cap = KQED_NOALLOC_PASS;
  w.ncap = KQED_NOALLOC_MAX_PAIRS;
  w.nbuckets = nb;
  w.groups = 1;
  w.rec = (double *)(d->noalloc + NA_REC);
  w.kinds = (unsigned *)(d->noalloc + NA_KINDS);
  w.key = (unsigned *)(d->noalloc + NA_KEY);
  w.rank = (unsigned *)(d->noalloc + NA_RANK);
  w.perm = (unsigned *)(d->noalloc + NA_PERM);
  w.count = (unsigned *)(d->noalloc + NA_COUNT);
  std::lock_guard<std::mutex> lk(d->mu);
  const kqed_tune keep = d->tune;
  const bool prof = d->profiling;
  if (d->tune.chunk > (size_t)KQED_NOALLOC_PASS) d->tune.chunk = KQED_NOALLOC_PASS;
  d->profiling = false; /* no timing events */
  const int rc = run_pairs(d, w, (hipStream_t)stream, which, M, n, x_dev, 4, y_dev, 4, out_dev, layout, nullptr, nullptr);
  d->tune = keep;
  d->profiling = prof;
  return rc;
}

extern "C" int kqed_get_tune(const struct QED_kernel_temps *t, kqed_tune *tune)
{
  if (!t || !t->dev || !tune) return -1;
  std::lock_guard<std::mutex> lk(t->dev->mu);
  *tune = t->dev->tune;
  return 0;
}

extern "C" int kqed_set_tune(const struct QED_kernel_temps *t, const kqed_tune *tune)
{
  if (!t || !t->dev || !tune) return -1;
  if (tune->chunk < 256 || tune->chunk > ((size_t)1 << 24) || tune->asm_grid < 1 || tune->asm_grid > (1 << 20) ||
      tune->sort_mode < KQED_SORT_CELL || tune->sort_mode > KQED_SORT_NONE) {
    set_err("kqed_set_tune: parameter out of range");
    return -1;
  }
  std::lock_guard<std::mutex> lk(t->dev->mu);
  t->dev->tune = *tune;
  return 0;
}

/* ---------------------------------------------------------------- ragged groups */
static int ragged_prepare(RaggedBufs &rb, size_t ng, const size_t *offsets, hipStream_t s)
{
  const size_t n = offsets[ng];
  for (size_t g = 0; g < ng; g++)
    if (offsets[g + 1] < offsets[g]) { set_err("ragged offsets must be non-decreasing"); return -1; }
  if (offsets[0] != 0) { set_err("ragged offsets must start at 0"); return -1; }
  if (n > 0xFFFFFFFFull || ng > 0x7FFFFFFF) { set_err("ragged batch too large"); return -1; }
  if (rb.ng < ng + 1 || rb.n < n) {
    /* the stream's previous ragged batch may still read the buffers being replaced */
    if (rb.off && hipStreamSynchronize(s) != hipSuccess) { set_err("stream synchronisation failed"); return -3; }
    rb.release();
    if (hipMalloc(&rb.off, 8 * (ng + 1)) != hipSuccess || (n && hipMalloc(&rb.ypair, 32 * n) != hipSuccess)) {
      set_err("ragged workspace allocation failed");
      return -2;
    }
    rb.ng = ng + 1;
    rb.n = n;
  }
  std::vector<unsigned long long> o(offsets, offsets + ng + 1);
  if (hipMemcpyAsync(rb.off, o.data(), 8 * (ng + 1), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    set_err("ragged offsets upload failed");
    return -3;
  }
  return 0;
}

static int ragged_launch(kqed_device *d, Workspace &w, RaggedBufs &rb, hipStream_t s, int which, double M, size_t ng,
                         size_t n, const double *y_dev, const double *x_dev, void *out_dev, int layout)
{
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_ragged_expand, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, (int)ng, rb.off, y_dev,
                     rb.ypair);
  return run_pairs(d, w, s, which, M, n, x_dev, 4, rb.ypair, 4, out_dev, layout, nullptr, nullptr);
}

extern "C" int kqed_batch_ragged(const struct QED_kernel_temps *t, int which, double M, size_t ngroups,
                                 const size_t *offsets, const double *y_dev, const double *x_dev, void *out_dev,
                                 int layout, void *stream)
{
  if (!t || !t->dev) { set_err("kqed_batch_ragged: tables not initialised"); return -1; }
  if (which < 0 || which > 4 || !layout_ok(layout) || !offsets) { set_err("kqed_batch_ragged: bad argument"); return -1; }
  if (ngroups == 0 || offsets[ngroups] == 0) return 0;
  kqed_device *d = t->dev;
  hipStream_t s = (hipStream_t)stream;
  /* the offsets and expanded y vectors live in the stream's workspace, like the pass buffers, so
   * batches enqueued on different streams never share them */
  std::lock_guard<std::mutex> lk(d->mu);
  Workspace &w = d->ws[stream];
  const int rc = ragged_prepare(w.rb, ngroups, offsets, s);
  if (rc) return rc;
  return ragged_launch(d, w, w.rb, s, which, M, ngroups, offsets[ngroups], y_dev, x_dev, out_dev, layout);
}

/* A ragged batch captured once into a HIP graph with its own workspace; every launch replays the
 * whole pipeline (expansion, sort, all passes) as one graph launch. */
struct kqed_graph {
  kqed_device *d = nullptr;
  Workspace w;
  RaggedBufs rb;
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  size_t n = 0;
};

extern "C" void kqed_graph_destroy(kqed_graph *g)
{
  if (!g) return;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->cap) { (void)hipStreamSynchronize(g->cap); (void)hipStreamDestroy(g->cap); }
  free_ws(g->w);
  g->rb.release();
  delete g;
}

extern "C" int kqed_graph_create_ragged(const struct QED_kernel_temps *t, int which, double M, size_t ngroups,
                                        const size_t *offsets, const double *y_dev, const double *x_dev, void *out_dev,
                                        int layout, kqed_graph **out)
{
  if (!out) return -1;
  *out = nullptr;
  if (!t || !t->dev) { set_err("kqed_graph_create_ragged: tables not initialised"); return -1; }
  if (which < 0 || which > 4 || !layout_ok(layout) || !offsets || ngroups == 0) {
    set_err("kqed_graph_create_ragged: bad argument");
    return -1;
  }
  kqed_device *d = t->dev;
  kqed_graph *g = new kqed_graph();
  g->d = d;
  g->n = offsets[ngroups];
  int rc = 0;
  if (hipStreamCreateWithFlags(&g->cap, hipStreamNonBlocking) != hipSuccess) { set_err("stream creation failed"); rc = -2; }
  if (!rc) rc = ragged_prepare(g->rb, ngroups, offsets, g->cap);
  if (!rc && g->n) {
    std::lock_guard<std::mutex> lk(d->mu);
    const size_t CH = d->tune.chunk;
    const bool sort = g->n >= d->tune.sort_min && d->tune.sort_mode != KQED_SORT_NONE;
    rc = ensure_ws(d, g->w, g->n < CH ? g->n : CH, sort ? g->n : 0);
    const bool prof = d->profiling;
    d->profiling = false; /* no event records inside the graph */
    if (!rc && hipStreamBeginCapture(g->cap, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      set_err("graph capture failed to start");
      rc = -3;
    } else if (!rc) {
      rc = ragged_launch(d, g->w, g->rb, g->cap, which, M, ngroups, g->n, y_dev, x_dev, out_dev, layout);
      const hipError_t e = hipStreamEndCapture(g->cap, &g->graph);
      if (!rc && e != hipSuccess) { set_err("graph capture failed: %s", hipGetErrorString(e)); rc = -3; }
      if (!rc && hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess) {
        set_err("graph instantiation failed");
        rc = -3;
      }
    }
    d->profiling = prof;
  }
  if (rc) { kqed_graph_destroy(g); return rc; }
  *out = g;
  return 0;
}

extern "C" int kqed_graph_launch(kqed_graph *g, void *stream)
{
  if (!g) { set_err("kqed_graph_launch: null graph"); return -1; }
  if (!g->exec) return 0; /* empty batch */
  const hipError_t e = hipGraphLaunch(g->exec, (hipStream_t)stream);
  if (e != hipSuccess) { set_err("graph launch failed: %s", hipGetErrorString(e)); return -3; }
  return 0;
}

/* ---------------------------------------------------------------- autotuner */
/* Times every candidate launch configuration on the caller's batch (device buffers, which/layout as
 * given) and keeps the fastest: passes of 2^18..2^22 pairs (no longer than the first pass that
 * covers the batch), assembly grids of 512, 1024 or 4096 workgroups, sort by cell or by x node.
 * Each candidate runs once untimed, then `reps` times. */
extern "C" int kqed_autotune(const struct QED_kernel_temps *t, int which, double M, size_t n, const double *x_dev,
                             const double *y_dev, void *out_dev, int layout, int reps, void *stream, kqed_tune *best,
                             double *best_ms)
{
  if (!t || !t->dev || !best) { set_err("kqed_autotune: bad argument"); return -1; }
  if (which < 0 || which > 4 || !layout_ok(layout) || n == 0) { set_err("kqed_autotune: bad argument"); return -1; }
  kqed_device *d = t->dev;
  hipStream_t s = (hipStream_t)stream;
  reps = reps < 1 ? 1 : reps;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) { set_err("event creation failed"); return -2; }
  kqed_tune orig;
  {
    std::lock_guard<std::mutex> lk(d->mu);
    orig = d->tune;
  }
  kqed_tune bt = orig;
  double bms = 1e300;
  int rc = 0;
  const size_t chunks[] = { (size_t)1 << 18, (size_t)1 << 19, (size_t)1 << 20, (size_t)1 << 21, (size_t)1 << 22 };
  const int grids[] = { 512, 1024, 4096 };
  const int modes[] = { KQED_SORT_CELL, KQED_SORT_XNODE };
  for (size_t ci = 0; ci < 5 && !rc; ci++) {
    if (ci > 0 && chunks[ci - 1] >= n) break; /* longer passes than the batch change nothing */
    for (int gi = 0; gi < 3 && !rc; gi++)
      for (int mi = 0; mi < 2 && !rc; mi++) {
        kqed_tune c = orig;
        c.chunk = chunks[ci];
        c.asm_grid = grids[gi];
        c.sort_mode = modes[mi];
        std::lock_guard<std::mutex> lk(d->mu);
        d->tune = c;
        Workspace &w = d->ws[stream];
        rc = run_pairs(d, w, s, which, M, n, x_dev, 4, y_dev, 4, out_dev, layout, nullptr, nullptr);
        if (rc) break;
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < reps && !rc; r++)
          rc = run_pairs(d, w, s, which, M, n, x_dev, 4, y_dev, 4, out_dev, layout, nullptr, nullptr);
        (void)hipEventRecord(e1, s);
        float ms = 0.f;
        if (!rc && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)) {
          set_err("autotune timing failed");
          rc = -3;
        }
        if (!rc && ms / reps < bms) { bms = ms / reps; bt = c; }
      }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::lock_guard<std::mutex> lk(d->mu);
  d->tune = rc ? orig : bt;
  if (rc) return rc;
  *best = bt;
  if (best_ms) *best_ms = bms;
  return 0;
}

/* contraction buffers: dots [nx], W [nx][KQED_WCOLS] (shortcut only) and the reduction scratch `red`:
 * block sums [RED_BLOCKS] | column partials [RED_BLOCKS][384] | sum_x T [384] | consts [2] |
 * zero 4-vector [4] | y [ny][4] | y-only kernels [ny][384] | sigma [ny] | result [ny] */
static const size_t RED_BLOCKS = 1024;
static size_t red_len(size_t ny) { return RED_BLOCKS + 384 * RED_BLOCKS + 384 + 2 + 4 + ny * (4 + 384 + 2); }

static int ensure_contract_ws(Workspace &w, size_t nx, size_t ny, bool need_W)
{
  if (w.xcap < nx || w.ycap < ny || (need_W && !w.W)) {
    (void)hipFree(w.dots); (void)hipFree(w.W); (void)hipFree(w.red);
    w.dots = w.W = w.red = nullptr;
    w.xcap = w.ycap = 0;
    if (hipMalloc(&w.dots, 8 * nx) != hipSuccess || hipMalloc(&w.red, 8 * red_len(ny)) != hipSuccess ||
        (need_W && hipMalloc(&w.W, 8 * (size_t)KQED_WCOLS * nx) != hipSuccess)) {
      set_err("contraction workspace allocation (%zu sites) failed", nx);
      return -2;
    }
    w.xcap = nx;
    w.ycap = ny;
  }
  return 0;
}

/* deterministic sum of in[0, n) into *out (two fixed-order passes; tmp holds RED_BLOCKS doubles) */
static void dev_sum(hipStream_t s, size_t n, const double *in, double *tmp, double *out)
{
  const size_t nb = std::min(RED_BLOCKS, (n + 255) / 256);
  const size_t per = (n + nb - 1) / nb;
  hipLaunchKernelGGL(k_block_sum, dim3((unsigned)nb), dim3(256), 0, s, n, per, in, tmp);
  hipLaunchKernelGGL(k_block_sum, dim3(1), dim3(256), 0, s, nb, nb, tmp, out);
}

extern "C" int kqed_contract(const struct QED_kernel_temps *t, int which, double M, size_t nx, const double *x_dev,
                             const float *T_dev, size_t ny, const double *y, double *s_out, void *stream)
{
  if (!t || !t->dev) { set_err("kqed_contract: tables not initialised"); return -1; }
  if (which < 0 || which > 4 || (ny && (!y || !s_out)) || (nx && (!x_dev || !T_dev))) {
    set_err("kqed_contract: bad argument");
    return -1;
  }
  if (nx > 0xFFFFFFFFull || ny > 65536) { set_err("kqed_contract: at most 2^32-1 sites and 65536 y points"); return -1; }
  if (ny == 0) return 0;
  if (nx == 0) { for (size_t j = 0; j < ny; j++) s_out[j] = 0.0; return 0; }
  kqed_device *d = t->dev;
  hipStream_t s = (hipStream_t)stream;
  std::lock_guard<std::mutex> lk(d->mu);
  Workspace &w = d->ws[stream];
  /* L0..L3 take the linear shortcut: sum_x L0(x,y).T(x) through the pulled-back site tensors
   * W = A^T T, and the subtraction terms factor into a y-only kernel contracted with sum_x T(x)
   * and an x-only sum computed once per call.  M-L2 contracts every pair in full. */
  const bool shortcut = which != KQED_ML2;
  if (ensure_contract_ws(w, nx, ny, shortcut)) return -2;
  double *part = w.red, *colpart = part + RED_BLOCKS, *sumT = colpart + 384 * RED_BLOCKS, *consts = sumT + 384,
         *zero4 = consts + 2, *yd = zero4 + 4, *La = yd + 4 * ny, *sigma = La + 384 * ny, *sres = sigma + ny;
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, (unsigned *)consts, (size_t)12, 0u);
  hipError_t e = hipMemcpyAsync(yd, y, 8 * 4 * ny, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) { set_err("kqed_contract: copy failed: %s", hipGetErrorString(e)); return -3; }
  int rc = 0;
  /* pull-back of the site tensors; for L1..L3 the same pass sums them per column (sum_x T) */
  const size_t cw_blocks = std::min((nx + 3) / 4, RED_BLOCKS);
  if (shortcut)
    hipLaunchKernelGGL(k_contract_weights, dim3((unsigned)cw_blocks), dim3(256), 0, s, nx, T_dev, d->csc_ent, d->csc_colx,
                       w.W, which != KQED_L0 ? colpart : nullptr);
  for (size_t j = 0; j < ny && !rc; j++) {
    rc = run_pairs(d, w, s, shortcut ? KQED_L0 : which, M, nx, x_dev, 4, yd + 4 * j, 0, w.dots, LAYOUT_DOT, T_dev,
                   shortcut ? w.W : nullptr);
    if (!rc) dev_sum(s, nx, w.dots, part, sigma + j);
  }
  double ca = 0.0, cc0 = 0.0, cc1 = 0.0;
  bool with_y = false;
  if (!rc && shortcut && which != KQED_L0) {
    hipLaunchKernelGGL(k_colsum_final, dim3(384), dim3(256), 0, s, (int)cw_blocks, colpart, sumT);
    with_y = true;
    /* x-only sums over the sites: consts[k] = sum_x L0(a_k(x), b_k(x)) . T(x) */
    const double *ax[2] = { nullptr, nullptr }, *bx[2] = { nullptr, nullptr };
    int as[2] = { 4, 4 }, bs[2] = { 4, 4 };
    if (which == KQED_L1) {
      /* L1 = L0(x,y) - L0(x,x)/2 - L0(y,y)/2 (reference kernels.c:119-135) */
      ax[0] = x_dev; bx[0] = x_dev; cc0 = -0.5; ca = -0.5;
      rc = run_pairs(d, w, s, KQED_L0, 0.0, ny, yd, 4, yd, 4, La, KQED_AOS, nullptr, nullptr);
    } else if (which == KQED_L2) {
      /* L2 = L0(x,y) - L0(0,y) - L0(x,0) (kernels.c:138-155) */
      ax[0] = x_dev; bx[0] = zero4; bs[0] = 0; cc0 = -1.0; ca = -1.0;
      rc = run_pairs(d, w, s, KQED_L0, 0.0, ny, zero4, 0, yd, 4, La, KQED_AOS, nullptr, nullptr);
    } else {
      /* L3 = L0(x,y) - L0(x,x) + L0(0,x) - L0(0,y) (kernels.c:158-177) */
      ax[0] = x_dev; bx[0] = x_dev; cc0 = -1.0;
      ax[1] = zero4; as[1] = 0; bx[1] = x_dev; cc1 = 1.0; ca = -1.0;
      rc = run_pairs(d, w, s, KQED_L0, 0.0, ny, zero4, 0, yd, 4, La, KQED_AOS, nullptr, nullptr);
    }
    for (int k = 0; k < 2 && !rc; k++) {
      if (!ax[k]) continue;
      rc = run_pairs(d, w, s, KQED_L0, 0.0, nx, ax[k], as[k], bx[k], bs[k], w.dots, LAYOUT_DOT, T_dev, nullptr);
      if (!rc) dev_sum(s, nx, w.dots, part, consts + k);
    }
  }
  if (rc) return rc;
  hipLaunchKernelGGL(k_combine, dim3((unsigned)ny), dim3(64), 0, s, (int)ny, sigma, with_y ? La : nullptr, ca, sumT,
                     with_y ? consts : nullptr, cc0, cc1, sres);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(s_out, sres, 8 * ny, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) { set_err("kqed_contract: %s", hipGetErrorString(e)); return -3; }
  return 0;
}

struct HostStage {
  size_t cap = 0;
  double *x = nullptr, *y = nullptr, *o = nullptr;
  hipStream_t s = nullptr;
  ~HostStage()
  {
    (void)hipFree(x); (void)hipFree(y); (void)hipFree(o);
    if (s) (void)hipStreamDestroy(s);
  }
};
static thread_local HostStage g_stage;

static size_t out_bytes(int layout) { return layout == KQED_AOS_F32 || layout == KQED_SOA_F32 ? 1536 : 3072; }

extern "C" int kqed_batch_host(const struct QED_kernel_temps *t, int which, double M, size_t n, const double *x,
                               const double *y, void *out, int layout)
{
  if (!t || !t->dev) { set_err("kqed_batch_host: tables not initialised"); return -1; }
  if (n == 0) return 0;
  HostStage &st = g_stage;
  if (!st.s && hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking) != hipSuccess) { set_err("stream creation failed"); return -2; }
  if (st.cap < n) {
    (void)hipFree(st.x); (void)hipFree(st.y); (void)hipFree(st.o);
    st.x = st.y = st.o = nullptr;
    st.cap = 0;
    if (hipMalloc(&st.x, 32 * n) != hipSuccess || hipMalloc(&st.y, 32 * n) != hipSuccess ||
        hipMalloc(&st.o, 3072 * n) != hipSuccess) {
      set_err("staging allocation failed");
      return -2;
    }
    st.cap = n;
  }
  if (hipMemcpyAsync(st.x, x, 32 * n, hipMemcpyHostToDevice, st.s) != hipSuccess ||
      hipMemcpyAsync(st.y, y, 32 * n, hipMemcpyHostToDevice, st.s) != hipSuccess) {
    set_err("host to device copy failed");
    return -3;
  }
  int rc = kqed_batch(t, which, M, n, st.x, st.y, st.o, layout, st.s);
  if (rc) return rc;
  if (hipMemcpyAsync(out, st.o, out_bytes(layout) * n, hipMemcpyDeviceToHost, st.s) != hipSuccess ||
      hipStreamSynchronize(st.s) != hipSuccess) {
    set_err("device to host copy failed");
    return -3;
  }
  return 0;
}

/* ------------------------------------------------------------------ kernel combinations */

/* derived argument pairs and L0 tensors of one combination pass (ne evaluations) */
static int ensure_comb_ws(Workspace &w, size_t ne)
{
  if (w.ccap >= ne) return 0;
  (void)hipFree(w.cargs); (void)hipFree(w.cL);
  w.cargs = w.cL = nullptr; w.ccap = 0;
  if (hipMalloc(&w.cargs, 64 * ne) != hipSuccess || hipMalloc(&w.cL, 3072 * ne) != hipSuccess) {
    set_err("combination workspace allocation of %zu evaluations failed", ne);
    return -2;
  }
  w.ccap = ne;
  return 0;
}

/* Passes of cp pairs (E * cp <= tune.chunk L0 evaluations): derived arguments, the L0 pipeline
 * over all of them (one sort per pass: the special x=0 / y=0 evaluations group together), then
 * the per-pair combination into the caller's output. */
extern "C" int kqed_combine(const struct QED_kernel_temps *t, int comb, const double *M, size_t n, const double *x_dev,
                            const double *y_dev, double *out_dev, void *stream)
{
  if (!t || !t->dev) { set_err("kqed_combine: tables not initialised"); return -1; }
  const int E = kqed_comb_evals(comb);
  if (!E) { set_err("kqed_combine: unknown combination %d", comb); return -1; }
  if (n == 0) return 0;
  if (n > 0xFFFFFFFFull) { set_err("kqed_combine: at most 2^32-1 pairs per call"); return -1; }
  kqed_device *d = t->dev;
  std::lock_guard<std::mutex> lk(d->mu);
  Workspace &w = d->ws[stream];
  hipStream_t s = (hipStream_t)stream;
  size_t cp = d->tune.chunk / (size_t)E;
  if (cp < 1) cp = 1;
  if (cp > n) cp = n;
  if (ensure_comb_ws(w, cp * E)) return -2;
  const size_t len = 384 * (size_t)kqed_comb_kernels(comb);
  for (size_t c0 = 0; c0 < n; c0 += cp) {
    const size_t c = n - c0 < cp ? n - c0 : cp;
    double *X = w.cargs, *Y = w.cargs + 4 * (size_t)E * c;
    kqed_comb_launch_args(comb, c, x_dev + 4 * c0, y_dev + 4 * c0, 4, X, Y, s);
    const int rc = run_pairs(d, w, s, KQED_L0, 0.0, (size_t)E * c, X, 4, Y, 4, w.cL, KQED_AOS, nullptr, nullptr);
    if (rc) return rc;
    kqed_comb_launch(comb, c, w.cL, x_dev + 4 * c0, y_dev + 4 * c0, 4, M, out_dev + len * c0, nullptr, 0, s);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) { set_err("kernel launch failed: %s", hipGetErrorString(e)); return -3; }
  return 0;
}

/* s[j][q] = sum_x sum_c K_q(x, y_j)_c T(x)_c for every output kernel q of a combination: per y and
 * pass, derived arguments, the L0 pipeline, then the combination kernel in its dot mode (the outputs
 * never leave LDS); per-pair dots summed over x in a fixed order. */
extern "C" int kqed_contract_comb(const struct QED_kernel_temps *t, int comb, const double *M, size_t nx,
                                  const double *x_dev, const float *T_dev, size_t ny, const double *y_host,
                                  double *s_host, void *stream)
{
  if (!t || !t->dev) { set_err("kqed_contract_comb: tables not initialised"); return -1; }
  const int E = kqed_comb_evals(comb), NK = kqed_comb_kernels(comb);
  if (!E) { set_err("kqed_contract_comb: unknown combination %d", comb); return -1; }
  if (ny == 0) return 0;
  if (nx > 0xFFFFFFFFull || ny > 65536) { set_err("kqed_contract_comb: at most 2^32-1 sites and 65536 y points"); return -1; }
  if (nx == 0) {
    memset(s_host, 0, 8 * (size_t)NK * ny);
    return 0;
  }
  kqed_device *d = t->dev;
  std::lock_guard<std::mutex> lk(d->mu);
  Workspace &w = d->ws[stream];
  hipStream_t s = (hipStream_t)stream;
  size_t cp = d->tune.chunk / (size_t)E;
  if (cp < 1) cp = 1;
  if (cp > nx) cp = nx;
  if (ensure_comb_ws(w, cp * E)) return -2;
  const size_t need = (size_t)NK * nx + (size_t)NK * ny + 4 * ny + RED_BLOCKS;
  if (w.cdcap < need) {
    (void)hipFree(w.cdots); w.cdots = nullptr; w.cdcap = 0;
    if (hipMalloc(&w.cdots, 8 * need) != hipSuccess) { set_err("contraction workspace allocation failed"); return -2; }
    w.cdcap = need;
  }
  double *dots = w.cdots, *sd = dots + (size_t)NK * nx, *yd = sd + (size_t)NK * ny, *tmp = yd + 4 * ny;
  if (hipMemcpyAsync(yd, y_host, 32 * ny, hipMemcpyHostToDevice, s) != hipSuccess) { set_err("y upload failed"); return -3; }
  for (size_t j = 0; j < ny; j++) {
    for (size_t c0 = 0; c0 < nx; c0 += cp) {
      const size_t c = nx - c0 < cp ? nx - c0 : cp;
      double *X = w.cargs, *Y = w.cargs + 4 * (size_t)E * c;
      kqed_comb_launch_args(comb, c, x_dev + 4 * c0, yd + 4 * j, 0, X, Y, s);
      const int rc = run_pairs(d, w, s, KQED_L0, 0.0, (size_t)E * c, X, 4, Y, 4, w.cL, KQED_AOS, nullptr, nullptr);
      if (rc) return rc;
      kqed_comb_launch(comb, c, w.cL, x_dev + 4 * c0, yd + 4 * j, 0, M, dots + c0, T_dev + 384 * c0, nx, s);
    }
    for (int q = 0; q < NK; q++) dev_sum(s, nx, dots + (size_t)q * nx, tmp, sd + j * NK + q);
  }
  if (hipMemcpyAsync(s_host, sd, 8 * (size_t)NK * ny, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    set_err("kqed_contract_comb: %s", hipGetErrorString(hipGetLastError()));
    return -3;
  }
  return 0;
}

struct CombStage {
  size_t cap = 0, ocap = 0;
  double *x = nullptr, *o = nullptr;
  hipStream_t s = nullptr;
  ~CombStage()
  {
    (void)hipFree(x); (void)hipFree(o);
    if (s) (void)hipStreamDestroy(s);
  }
};
static thread_local CombStage g_cstage;

extern "C" int kqed_combine_host(const struct QED_kernel_temps *t, int comb, const double *M, size_t n, const double *x,
                                 const double *y, double *out)
{
  if (!t || !t->dev) { set_err("kqed_combine_host: tables not initialised"); return -1; }
  const size_t len = 384 * (size_t)kqed_comb_kernels(comb);
  if (!len) { set_err("kqed_combine_host: unknown combination %d", comb); return -1; }
  if (n == 0) return 0;
  CombStage &st = g_cstage;
  if (!st.s && hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking) != hipSuccess) { set_err("stream creation failed"); return -2; }
  if (st.cap < n) {
    (void)hipFree(st.x); st.x = nullptr; st.cap = 0;
    if (hipMalloc(&st.x, 64 * n) != hipSuccess) { set_err("staging allocation failed"); return -2; }
    st.cap = n;
  }
  if (st.ocap < len * n) {
    (void)hipFree(st.o); st.o = nullptr; st.ocap = 0;
    if (hipMalloc(&st.o, 8 * len * n) != hipSuccess) { set_err("staging allocation failed"); return -2; }
    st.ocap = len * n;
  }
  double *xd = st.x, *yd = st.x + 4 * n;
  if (hipMemcpyAsync(xd, x, 32 * n, hipMemcpyHostToDevice, st.s) != hipSuccess ||
      hipMemcpyAsync(yd, y, 32 * n, hipMemcpyHostToDevice, st.s) != hipSuccess) {
    set_err("host to device copy failed");
    return -3;
  }
  const int rc = kqed_combine(t, comb, M, n, xd, yd, st.o, st.s);
  if (rc) return rc;
  if (hipMemcpyAsync(out, st.o, 8 * len * n, hipMemcpyDeviceToHost, st.s) != hipSuccess ||
      hipStreamSynchronize(st.s) != hipSuccess) {
    set_err("device to host copy failed");
    return -3;
  }
  return 0;
}

/* ------------------------------------------------------------------ perturbative lepton loop */

static int ensure_pert_tab(kqed_device *d, const struct QED_kernel_temps *t, hipStream_t s)
{
  if (d->pert_tab) return 0;
  if (!t->G8) { set_err("lepton loop: the G8 trace table is missing"); return -1; }
  std::vector<unsigned> h(kqed_pert_table_bytes() / 4);
  if (kqed_pert_build_table(t->G8, h.data())) { set_err("lepton loop: the G8 table is not a table of gamma traces"); return -1; }
  unsigned *p = nullptr;
  if (hipMalloc(&p, 4 * h.size()) != hipSuccess) { set_err("lepton loop: table allocation failed"); return -2; }
  if (hipMemcpyAsync(p, h.data(), 4 * h.size(), hipMemcpyHostToDevice, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
    (void)hipFree(p);
    set_err("lepton loop: table upload failed");
    return -3;
  }
  d->pert_tab = p;
  return 0;
}

static int ensure_pert_ws(Workspace &w, size_t c)
{
  if (w.pcap >= c) return 0;
  (void)hipFree(w.pP); (void)hipFree(w.pX);
  w.pP = w.pX = nullptr; w.pcap = 0;
  if (hipMalloc(&w.pP, 8 * 384 * c) != hipSuccess || hipMalloc(&w.pX, 64 * c) != hipSuccess) {
    set_err("lepton loop workspace allocation of %zu pairs failed", c);
    return -2;
  }
  w.pcap = c;
  return 0;
}

/* pairs per lepton-loop pass: 3 KiB of iPihat and 64 B of scaled pairs each */
static const size_t PERT_CHUNK = (size_t)1 << 18;

extern "C" int kqed_pihat(const struct QED_kernel_temps *t, size_t n, const double *x_dev, const double *y_dev,
                          double *out_dev, void *stream)
{
  if (!t || !t->dev) { set_err("kqed_pihat: tables not initialised"); return -1; }
  if (n == 0) return 0;
  kqed_device *d = t->dev;
  std::lock_guard<std::mutex> lk(d->mu);
  hipStream_t s = (hipStream_t)stream;
  const int rc = ensure_pert_tab(d, t, s);
  if (rc) return rc;
  kqed_pert_launch_ipihat(n, x_dev, y_dev, d->pert_tab, out_dev, s);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) { set_err("kernel launch failed: %s", hipGetErrorString(e)); return -3; }
  return 0;
}

extern "C" int kqed_lepton_loop(const struct QED_kernel_temps *t, int which, double M, double Mv, size_t n,
                                const double *x_dev, const double *y_dev, double *out_dev, void *stream)
{
  if (!t || !t->dev) { set_err("kqed_lepton_loop: tables not initialised"); return -1; }
  if (which < 0 || which > 4) { set_err("kqed_lepton_loop: bad kernel"); return -1; }
  if (n == 0) return 0;
  if (n > 0xFFFFFFFFull / 3) { set_err("kqed_lepton_loop: too many pairs per call"); return -1; }
  kqed_device *d = t->dev;
  std::lock_guard<std::mutex> lk(d->mu);
  hipStream_t s = (hipStream_t)stream;
  int rc = ensure_pert_tab(d, t, s);
  if (rc) return rc;
  Workspace &w = d->ws[stream];
  const size_t CH = n < PERT_CHUNK ? n : PERT_CHUNK;
  if ((rc = ensure_pert_ws(w, CH))) return rc;
  for (size_t c0 = 0; c0 < n; c0 += CH) {
    const size_t c = n - c0 < CH ? n - c0 : CH;
    double *xs = w.pX, *ys = w.pX + 4 * c;
    kqed_pert_launch_scale(c, Mv, x_dev + 4 * c0, y_dev + 4 * c0, xs, ys, s);
    kqed_pert_launch_ipihat(c, xs, ys, d->pert_tab, w.pP, s);
    rc = run_pairs(d, w, s, which, M, c, x_dev + 4 * c0, 4, y_dev + 4 * c0, 4, out_dev + c0, LAYOUT_DOT64, w.pP, nullptr);
    if (rc) return rc;
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) { set_err("kernel launch failed: %s", hipGetErrorString(e)); return -3; }
  return 0;
}

/* one pair on host buffers through a staging stream (the reference's per-pair routines) */
struct PertStage {
  double *b = nullptr;
  hipStream_t s = nullptr;
  ~PertStage()
  {
    (void)hipFree(b);
    if (s) (void)hipStreamDestroy(s);
  }
};
static thread_local PertStage g_pstage;

static int pert_one(const struct QED_kernel_temps *t, bool antisym, const double *x, const double *y, double *out)
{
  if (!t || !t->dev) { set_err("lepton loop: tables not initialised"); return -1; }
  PertStage &st = g_pstage;
  if (!st.s && hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking) != hipSuccess) { set_err("stream creation failed"); return -2; }
  if (!st.b && hipMalloc(&st.b, 8 * (8 + 1280 + 384)) != hipSuccess) { set_err("staging allocation failed"); return -2; }
  double *xd = st.b, *yd = st.b + 4, *o = st.b + 8;
  if (hipMemcpyAsync(xd, x, 32, hipMemcpyHostToDevice, st.s) != hipSuccess ||
      hipMemcpyAsync(yd, y, 32, hipMemcpyHostToDevice, st.s) != hipSuccess) {
    set_err("host to device copy failed");
    return -3;
  }
  int rc;
  size_t len;
  if (antisym) {
    rc = kqed_pihat(t, 1, xd, yd, o, st.s);
    len = 384;
  } else {
    kqed_device *d = t->dev;
    std::lock_guard<std::mutex> lk(d->mu);
    rc = ensure_pert_tab(d, t, st.s);
    if (!rc) kqed_pert_launch_pihat1(1, xd, yd, d->pert_tab, o, st.s);
    len = 1280;
  }
  if (rc) return rc;
  if (hipMemcpyAsync(out, o, 8 * len, hipMemcpyDeviceToHost, st.s) != hipSuccess || hipStreamSynchronize(st.s) != hipSuccess) {
    set_err("device to host copy failed");
    return -3;
  }
  return 0;
}

extern "C" int kqed_pihat_host(const struct QED_kernel_temps *t, size_t n, const double *x, const double *y, double *out)
{
  for (size_t i = 0; i < n; i++) {
    const int rc = pert_one(t, true, x + 4 * i, y + 4 * i, out + 384 * i);
    if (rc) return rc;
  }
  return 0;
}

extern "C" void ipihatFermLoop_antisym(const double x[4], const double y[4], const struct QED_kernel_temps t,
                                       double vpihat[6][4][4][4])
{
  if (pert_one(&t, true, x, y, &vpihat[0][0][0][0])) {
    fprintf(stderr, "[KQED] %s\n", g_err);
    memset(vpihat, 0, 384 * sizeof(double));
  }
}

extern "C" void pihat1(const double x[4], const double y[4], const struct QED_kernel_temps t,
                       double vpihat1[4][4][4][4][4], double vpihatr1[4][4][4][4])
{
  double o[1280];
  if (pert_one(&t, false, x, y, o)) {
    fprintf(stderr, "[KQED] %s\n", g_err);
    memset(o, 0, sizeof o);
  }
  memcpy(vpihat1, o, 1024 * sizeof(double));
  memcpy(vpihatr1, o + 1024, 256 * sizeof(double));
}

extern "C" void init_g8(struct QED_kernel_temps *t)
{
  if (t && t->G8) kqed_fill_g8(t->G8);
}

/* the reference's per-pair combination routines, one pair through kqed_combine_host */
static void comb_one(int comb, const double *M, const double *xv, const double *yv, const struct QED_kernel_temps *t, void *K)
{
  if (kqed_combine_host(t, comb, M, 1, xv, yv, (double *)K) != 0) {
    fprintf(stderr, "[KQED] %s\n", g_err);
    memset(K, 0, 8 * 384 * (size_t)kqed_comb_kernels(comb));
  }
}

extern "C" void compute_all_kernels(const double xv[4], const double yv[4], const struct QED_kernel_temps t, struct QED_Kernels *K)
{ comb_one(KQED_COMB_ALL, nullptr, xv, yv, &t, K); }
extern "C" void compute_all_Mkernels(const double M[4], const double xv[4], const double yv[4], const struct QED_kernel_temps t,
                                     struct QED_Kernels *K)
{ comb_one(KQED_COMB_ALL_M, M, xv, yv, &t, K); }
extern "C" void compute_all_Mkernels_v2(const double M[4], const double xv[4], const double yv[4],
                                        const struct QED_kernel_temps t, struct QED_Kernels *K)
{ comb_one(KQED_COMB_ALL_M_V2, M, xv, yv, &t, K); }
extern "C" void compute_con_kernels(const double xv[4], const double yv[4], const struct QED_kernel_temps t, struct QED_Kernels *K)
{ comb_one(KQED_COMB_CON, nullptr, xv, yv, &t, K); }
extern "C" void compute_con_kernels_v2(const double xv[4], const double yv[4], const struct QED_kernel_temps t, struct Kernels *K)
{ comb_one(KQED_COMB_CON_V2, nullptr, xv, yv, &t, K); }
extern "C" void compute_con_kernelsM_L2(const double M[4], const double xv[4], const double yv[4],
                                        const struct QED_kernel_temps t, struct Kernels *K)
{ comb_one(KQED_COMB_CON_M_L2, M, xv, yv, &t, K); }
extern "C" void compute_sub_kernelsM_L2(const double M[4], const double xv[4], const double yv[4],
                                        const struct QED_kernel_temps t, struct Kernels *K)
{ comb_one(KQED_COMB_SUB_M_L2, M, xv, yv, &t, K); }
extern "C" void compute_all_kernels_SYMXY(const double xv[4], const double yv[4], const struct QED_kernel_temps t,
                                          struct QED_Kernels *K)
{ comb_one(KQED_COMB_SYMXY, nullptr, xv, yv, &t, K); }
extern "C" void compute_all_kernels_SYMXY_v2(const double xv[4], const double yv[4], const struct QED_kernel_temps t,
                                             struct Kernels *K)
{ comb_one(KQED_COMB_SYMXY_V2, nullptr, xv, yv, &t, K); }
extern "C" void compute_all_kernels_SYMXY0(const double xv[4], const double yv[4], const struct QED_kernel_temps t,
                                           struct QED_Kernels *K)
{ comb_one(KQED_COMB_SYMXY0, nullptr, xv, yv, &t, K); }
extern "C" void compute_all_kernels_SYMXY0_v2(const double xv[4], const double yv[4], const struct QED_kernel_temps t,
                                              struct Kernels *K)
{ comb_one(KQED_COMB_SYMXY0_V2, nullptr, xv, yv, &t, K); }

/* K = (K + W) / 2 elementwise over nk kernels (all_kernels.c:18-82, con_kernel.c:13-59) */
static void average_kernels(double *K, const double *W, size_t nk)
{
  for (size_t i = 0; i < 384 * nk; i++) K[i] = 0.5 * (K[i] + W[i]);
}
extern "C" void average_all_kernels(struct Kernels *K, const struct Kernels W)
{ average_kernels(&K->L0[0][0][0][0], &W.L0[0][0][0][0], 4); }
extern "C" void average_all_QED_kernels(struct QED_Kernels *K, const struct QED_Kernels W)
{ average_kernels(&K->L0.xy[0][0][0][0], &W.L0.xy[0][0][0][0], 8); }
extern "C" void average_con_kernels(struct Kernels *K, const struct Kernels *W)
{ average_kernels(&K[0].L0[0][0][0][0], &W[0].L0[0][0][0][0], 12); }

/* mu <-> nu of every .yx kernel (all_kernels.c:372-406) */
extern "C" void swap_munu_Lyx(struct QED_Kernels *K)
{
  struct Kernel *ks[4] = { &K->L0, &K->L1, &K->L2, &K->L3 };
  for (struct Kernel *k : ks)
    for (int rs = 0; rs < 6; rs++)
      for (int mu = 0; mu < 4; mu++)
        for (int nu = mu + 1; nu < 4; nu++)
          for (int la = 0; la < 4; la++) std::swap(k->yx[rs][mu][nu][la], k->yx[rs][nu][mu][la]);
}

static void one(int which, double M, const double *xv, const double *yv, const struct QED_kernel_temps *t, double *k)
{
  if (kqed_batch_host(t, which, M, 1, xv, yv, k, KQED_AOS) != 0) {
    /* the reference has no error return here: report and leave an all-zero kernel */
    memset(k, 0, 384 * sizeof(double));
  }
}

extern "C" void QED_kernel_L0(const double xv[4], const double yv[4], const struct QED_kernel_temps t, double kerv[6][4][4][4])
{ one(KQED_L0, 0.0, xv, yv, &t, &kerv[0][0][0][0]); }
extern "C" void QED_kernel_L1(const double xv[4], const double yv[4], const struct QED_kernel_temps t, double kerv[6][4][4][4])
{ one(KQED_L1, 0.0, xv, yv, &t, &kerv[0][0][0][0]); }
extern "C" void QED_kernel_L2(const double xv[4], const double yv[4], const struct QED_kernel_temps t, double kerv[6][4][4][4])
{ one(KQED_L2, 0.0, xv, yv, &t, &kerv[0][0][0][0]); }
extern "C" void QED_kernel_L3(const double xv[4], const double yv[4], const struct QED_kernel_temps t, double kerv[6][4][4][4])
{ one(KQED_L3, 0.0, xv, yv, &t, &kerv[0][0][0][0]); }
extern "C" void QED_Mkernel_L2(const double M, const double xv[4], const double yv[4], const struct QED_kernel_temps t,
                               double kerv[6][4][4][4])
{ one(KQED_ML2, M, xv, yv, &t, &kerv[0][0][0][0]); }

extern "C" int kqed_write_synthetic_ff(const char *path, int nstpx, int nstpy, int nf0, int nf1, double xh, double yh,
                                       unsigned long long seed)
{
  kqed_synth_params p;
  p.nstpx = nstpx; p.nstpy = nstpy; p.nf0 = nf0; p.nf1 = nf1; p.xh = xh; p.yh = yh; p.seed = seed;
  return kqed_synth_ff(path, &p);
}

/* ================================================================ learned surrogate */
/* Parameter layout (FP32, contiguous): layer l = 0..4 holds W_l[K_l][N_l] then b_l[N_l], with
 * K = {16, 128, 128, 128, 128} (rows 3..15 of W_0 stay zero) and N = {128, 128, 128, 128, 48}
 * (columns 34..47 of W_4 and b_4 stay zero). */
static const int SK[5] = { 16, 128, 128, 128, 128 };
static const int SN[5] = { 128, 128, 128, 128, 48 };
static size_t surr_off_W(int l)
{
  size_t o = 0;
  for (int i = 0; i < l; i++) o += (size_t)SK[i] * SN[i] + SN[i];
  return o;
}
static size_t surr_off_b(int l) { return surr_off_W(l) + (size_t)SK[l] * SN[l]; }
static const size_t SURR_NPARAM = 2048 + 128 + 3 * (16384 + 128) + 6144 + 48;
static const int NBB = KQED_SURR_NB * KQED_SURR_NB;

struct kqed_surrogate {
  kqed_device *d = nullptr;
  hipStream_t s = nullptr;
  float *P = nullptr, *G = nullptr, *Am = nullptr, *Av = nullptr; /* params, grads, Adam moments */
  unsigned short *Wp = nullptr;                                    /* packed BF16 inference weights */
  float *lsc = nullptr;                                            /* [34][NB][NB] log output scales */
  unsigned char *mask = nullptr;                                   /* [NB*NB] error-check mask */
  std::vector<float> h_lsc;
  std::vector<unsigned char> h_mask;
  SurrRange R;
  bool validated = false;
  long steps = 0;
  /* training buffers for batches of up to bcap samples */
  int bcap = 0;
  float *X = nullptr, *Tg = nullptr, *Wt = nullptr, *Z[5] = {}, *A[5] = {}, *dA = nullptr, *dZ = nullptr,
        *part = nullptr, *lpart = nullptr;
  unsigned *idx = nullptr;
};

static size_t surr_wp_off(int l)
{
  size_t o = 0;
  for (int i = 0; i < l; i++) o += (size_t)SK[i] * SN[i];
  return o;
}

static SurrNet surr_net(const kqed_surrogate *g)
{
  SurrNet n;
  for (int l = 0; l < 5; l++) {
    n.W[l] = g->Wp + surr_wp_off(l);
    n.b[l] = g->P + surr_off_b(l);
  }
  n.lsc = g->lsc;
  n.R = g->R;
  return n;
}

static void surr_free_train(kqed_surrogate *g)
{
  (void)hipFree(g->X); (void)hipFree(g->Tg); (void)hipFree(g->Wt); (void)hipFree(g->dA); (void)hipFree(g->dZ);
  (void)hipFree(g->part); (void)hipFree(g->lpart); (void)hipFree(g->idx);
  for (int l = 0; l < 5; l++) { (void)hipFree(g->Z[l]); (void)hipFree(g->A[l]); g->Z[l] = g->A[l] = nullptr; }
  g->X = g->Tg = g->Wt = g->dA = g->dZ = g->part = g->lpart = nullptr;
  g->idx = nullptr;
  g->bcap = 0;
}

extern "C" void kqed_surrogate_destroy(kqed_surrogate *g)
{
  if (!g) return;
  if (g->s) (void)hipStreamSynchronize(g->s);
  surr_free_train(g);
  (void)hipFree(g->P); (void)hipFree(g->G); (void)hipFree(g->Am); (void)hipFree(g->Av); (void)hipFree(g->Wp);
  (void)hipFree(g->lsc); (void)hipFree(g->mask);
  if (g->s) (void)hipStreamDestroy(g->s);
  delete g;
}

/* splitmix64: seeded host initialisation */
static unsigned long long smix(unsigned long long &s)
{
  unsigned long long z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* BF16 packing of the weights for the inference kernel (fragment order, see kqed_surrogate.hip) */
static int surr_publish(kqed_surrogate *g)
{
  std::vector<float> P(SURR_NPARAM);
  if (hipMemcpy(P.data(), g->P, 4 * SURR_NPARAM, hipMemcpyDeviceToHost) != hipSuccess) return -3;
  std::vector<unsigned short> W(surr_wp_off(5));
  for (int l = 0; l < 5; l++) {
    const float *w = P.data() + surr_off_W(l);
    unsigned short *o = W.data() + surr_wp_off(l);
    for (int k = 0; k < SK[l]; k++)
      for (int n = 0; n < SN[l]; n++) {
        unsigned u;
        memcpy(&u, &w[k * SN[l] + n], 4);
        u += 0x7FFFu + ((u >> 16) & 1u);
        o[((size_t)(k / 4) * SN[l] + n) * 4 + (k % 4)] = (unsigned short)(u >> 16);
      }
  }
  return hipMemcpy(g->Wp, W.data(), 2 * W.size(), hipMemcpyHostToDevice) == hipSuccess ? 0 : -3;
}

extern "C" int kqed_surrogate_create(const struct QED_kernel_temps *t, unsigned long long seed, kqed_surrogate **out)
{
  if (!t || !t->dev || !out) { set_err("kqed_surrogate_create: bad argument"); return -1; }
  *out = nullptr;
  kqed_surrogate *g = new kqed_surrogate();
  g->d = t->dev;
  const size_t np = SURR_NPARAM;
  if (hipStreamCreateWithFlags(&g->s, hipStreamNonBlocking) != hipSuccess || hipMalloc(&g->P, 4 * np) != hipSuccess ||
      hipMalloc(&g->G, 4 * np) != hipSuccess || hipMalloc(&g->Am, 4 * np) != hipSuccess ||
      hipMalloc(&g->Av, 4 * np) != hipSuccess || hipMalloc(&g->Wp, 2 * surr_wp_off(5)) != hipSuccess ||
      hipMalloc(&g->lsc, 4 * (size_t)KQED_NSURR * NBB) != hipSuccess || hipMalloc(&g->mask, NBB) != hipSuccess) {
    set_err("surrogate allocation failed");
    kqed_surrogate_destroy(g);
    return -2;
  }
  /* uniform(-a, a) weights with a = sqrt(6 / fan_in) (fan_in 3 for the first layer), zero biases */
  std::vector<float> P(np, 0.f);
  unsigned long long st = seed;
  for (int l = 0; l < 5; l++) {
    const int fin = l == 0 ? 3 : SK[l], nout = l == 4 ? KQED_NSURR : SN[l];
    const float a = sqrtf(6.f / fin) * (l == 4 ? 0.5f : 1.f);
    float *w = P.data() + surr_off_W(l);
    for (int k = 0; k < fin; k++)
      for (int n = 0; n < nout; n++) w[k * SN[l] + n] = a * (2.f * (float)((smix(st) >> 11) * 0x1.0p-53) - 1.f);
  }
  const kqed_dev_tables &T = g->d->dt;
  double xx0, xx1, yy1;
  if (hipMemcpy(&xx0, T.XX, 8, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(&xx1, T.XX + T.nstpx - 1, 8, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(&yy1, T.YY + T.nstpy - 1, 8, hipMemcpyDeviceToHost) != hipSuccess) {
    set_err("surrogate: grid read failed");
    kqed_surrogate_destroy(g);
    return -3;
  }
  g->R = { (float)log(xx0), (float)log(xx1), (float)log(KQED_SURR_YMIN), (float)log(yy1) };
  g->h_lsc.assign((size_t)KQED_NSURR * NBB, 0.f);
  g->h_mask.assign(NBB, 0);
  if (hipMemcpy(g->P, P.data(), 4 * np, hipMemcpyHostToDevice) != hipSuccess || hipMemset(g->Am, 0, 4 * np) != hipSuccess ||
      hipMemset(g->Av, 0, 4 * np) != hipSuccess || hipMemset(g->G, 0, 4 * np) != hipSuccess ||
      hipMemcpy(g->lsc, g->h_lsc.data(), 4 * g->h_lsc.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(g->mask, 0, NBB) != hipSuccess || surr_publish(g)) {
    set_err("surrogate initialisation failed");
    kqed_surrogate_destroy(g);
    return -3;
  }
  *out = g;
  return 0;
}

static int surr_ensure_batch(kqed_surrogate *g, int B)
{
  if (g->bcap >= B) return 0;
  surr_free_train(g);
  bool ok = hipMalloc(&g->X, 4L * B * 16) == hipSuccess && hipMalloc(&g->Tg, 4L * B * KQED_NSURR) == hipSuccess &&
            hipMalloc(&g->Wt, 4L * B * KQED_NSURR) == hipSuccess && hipMalloc(&g->dA, 4L * B * 128) == hipSuccess &&
            hipMalloc(&g->dZ, 4L * B * 128) == hipSuccess && hipMalloc(&g->idx, 4L * B) == hipSuccess &&
            hipMalloc(&g->lpart, 4L * ((B * 48 + 255) / 256)) == hipSuccess &&
            hipMalloc(&g->part, 4L * 64 * 128 * 128) == hipSuccess;
  for (int l = 0; l < 5 && ok; l++)
    ok = hipMalloc(&g->Z[l], 4L * B * SN[l]) == hipSuccess && hipMalloc(&g->A[l], 4L * B * SN[l]) == hipSuccess;
  if (!ok) { surr_free_train(g); set_err("surrogate training buffers failed"); return -2; }
  g->bcap = B;
  return 0;
}

static void gemm(hipStream_t s, int M, int N, int K, const float *A, long sam, long sak, const float *B, long sbk, long sbn,
                 float *C, long ldc, const float *bias, float *part)
{
  /* split the reduction when the output grid alone cannot fill the device */
  const int tiles = ((M + 31) / 32) * ((N + 31) / 32);
  int splits = 1;
  while (splits < 64 && tiles * splits < 512 && K / (splits * 2) >= 256) splits *= 2;
  const int kper = ((K + splits - 1) / splits + 15) / 16 * 16;
  splits = (K + kper - 1) / kper;
  if (splits == 1) {
    hipLaunchKernelGGL(k_gemm_bf16, dim3((N + 31) / 32, (M + 31) / 32, 1), dim3(64), 0, s, M, N, K, kper, A, sam, sak, B, sbk,
                       sbn, C, ldc, 0L, bias);
    return;
  }
  hipLaunchKernelGGL(k_gemm_bf16, dim3((N + 31) / 32, (M + 31) / 32, splits), dim3(64), 0, s, M, N, K, kper, A, sam, sak, B,
                     sbk, sbn, part, (long)N, (long)M * N, bias);
  hipLaunchKernelGGL(k_sum_splits, dim3((unsigned)(((long)M * N + 255) / 256)), dim3(256), 0, s, (long)M * N, splits,
                     (long)M * N, part, C); /* ldc == N here */
}

/* Forward and backward of one batch already in X/Tg/Wt: gradients into G, loss partials */
static void surr_fwd_bwd(kqed_surrogate *g, int B)
{
  hipStream_t s = g->s;
  const long n128 = (long)B * 128;
  const float *in = g->X;
  int ldin = 16;
  for (int l = 0; l < 5; l++) {
    gemm(s, B, SN[l], SK[l], in, ldin, 1, g->P + surr_off_W(l), SN[l], 1, g->Z[l], SN[l], g->P + surr_off_b(l), g->part);
    if (l < 4) hipLaunchKernelGGL(k_silu, dim3((unsigned)((n128 + 255) / 256)), dim3(256), 0, s, n128, g->Z[l], g->A[l]);
    in = g->A[l];
    ldin = SN[l];
  }
  const long nout = (long)B * 48;
  hipLaunchKernelGGL(k_mse, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, s, B, 48, KQED_NSURR, g->Z[4], g->Tg, g->Wt,
                     (float)B * KQED_NSURR, g->dZ, g->lpart);
  for (int l = 4; l >= 0; l--) {
    /* dZ holds dL/dZ_l [B][N_l]; weight gradient A_{l-1}^T dZ_l, bias gradient column sums */
    const float *Ain = l ? g->A[l - 1] : g->X;
    const long ldA = l ? 128 : 16;
    gemm(s, SK[l], SN[l], B, Ain, 1, ldA, g->dZ, SN[l], 1, g->G + surr_off_W(l), SN[l], nullptr, g->part);
    hipLaunchKernelGGL(k_colsum_f, dim3((SN[l] + 63) / 64), dim3(1024), 0, s, B, SN[l], g->dZ, g->G + surr_off_b(l));
    if (l == 0) break;
    /* dA_{l-1} = dZ_l W_l^T, then dZ_{l-1} = dA_{l-1} silu'(Z_{l-1}) */
    gemm(s, B, SK[l], SN[l], g->dZ, SN[l], 1, g->P + surr_off_W(l), 1, SN[l], g->dA, SK[l], nullptr, g->part);
    hipLaunchKernelGGL(k_silu_bwd, dim3((unsigned)((n128 + 255) / 256)), dim3(256), 0, s, n128, g->Z[l - 1], g->dA, g->dZ);
  }
}

static double surr_loss(kqed_surrogate *g, int B)
{
  const int nb = (B * 48 + 255) / 256;
  std::vector<float> lp(nb);
  if (hipMemcpyAsync(lp.data(), g->lpart, 4L * nb, hipMemcpyDeviceToHost, g->s) != hipSuccess ||
      hipStreamSynchronize(g->s) != hipSuccess)
    return -1.0;
  double l = 0.0;
  for (float v : lp) l += v;
  return l / ((double)B * KQED_NSURR);
}

/* Training data: the FP64 interpolator on the given pairs (device), rows of KQED_SURR_ROW doubles */
static int surr_extract(kqed_surrogate *g, size_t n, const double *x_dev, const double *y_dev, double *data)
{
  kqed_device *d = g->d;
  std::lock_guard<std::mutex> lk(d->mu);
  return run_pairs(d, d->ws[(void *)g->s], g->s, KQED_L0, 0.0, n, x_dev, 4, y_dev, 4, data, LAYOUT_SCALARS, nullptr, nullptr);
}

/* per-bin log maxima of every weight function, empty bins filled from the nearest filled bin */
static int surr_scales(kqed_surrogate *g, size_t n, const double *data)
{
  unsigned *bm = nullptr, *bc = nullptr;
  const size_t nm = (size_t)KQED_NSURR * NBB;
  if (hipMalloc(&bm, 4 * nm) != hipSuccess || hipMalloc(&bc, 4 * NBB) != hipSuccess) {
    (void)hipFree(bm);
    set_err("surrogate scale buffers failed");
    return -2;
  }
  int rc = 0;
  std::vector<unsigned> hm(nm), hc(NBB);
  if (hipMemsetAsync(bm, 0, 4 * nm, g->s) != hipSuccess || hipMemsetAsync(bc, 0, 4 * NBB, g->s) != hipSuccess) rc = -3;
  if (!rc) {
    hipLaunchKernelGGL(k_surr_binmax, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, g->s, (long)n, data, g->R, bm, bc);
    if (hipMemcpyAsync(hm.data(), bm, 4 * nm, hipMemcpyDeviceToHost, g->s) != hipSuccess ||
        hipMemcpyAsync(hc.data(), bc, 4 * NBB, hipMemcpyDeviceToHost, g->s) != hipSuccess ||
        hipStreamSynchronize(g->s) != hipSuccess)
      rc = -3;
  }
  (void)hipFree(bm);
  (void)hipFree(bc);
  if (rc) { set_err("surrogate scale pass failed"); return rc; }
  for (int k = 0; k < KQED_NSURR; k++)
    for (int c = 0; c < NBB; c++) {
      int best = -1, bd = 1 << 30;
      for (int e = 0; e < NBB; e++) {
        if (!hm[(size_t)k * NBB + e]) continue;
        const int du = c / KQED_SURR_NB - e / KQED_SURR_NB, dv = c % KQED_SURR_NB - e % KQED_SURR_NB;
        const int dd = du * du + dv * dv;
        if (dd < bd) { bd = dd; best = e; }
        if (!dd) break;
      }
      float v = 1.f;
      if (best >= 0) memcpy(&v, &hm[(size_t)k * NBB + best], 4);
      g->h_lsc[(size_t)k * NBB + c] = logf(v > 1e-30f ? v : 1e-30f);
    }
  return hipMemcpy(g->lsc, g->h_lsc.data(), 4 * nm, hipMemcpyHostToDevice) == hipSuccess ? 0 : -3;
}

extern "C" int kqed_surrogate_train(kqed_surrogate *g, size_t n, const double *x_dev, const double *y_dev, int epochs,
                                    int batch, double lr, unsigned long long seed, double *loss_out)
{
  if (!g || !x_dev || !y_dev || n == 0 || epochs < 1 || batch < 64 || batch > (1 << 16)) {
    set_err("kqed_surrogate_train: bad argument");
    return -1;
  }
  double *data = nullptr;
  if (hipMalloc(&data, 8 * KQED_SURR_ROW * n) != hipSuccess) { set_err("surrogate data allocation failed"); return -2; }
  int rc = surr_extract(g, n, x_dev, y_dev, data);
  if (!rc) rc = surr_scales(g, n, data);
  if (!rc) rc = surr_ensure_batch(g, batch);
  std::vector<unsigned> hidx(batch);
  unsigned long long st = seed;
  const long steps = (long)epochs * (long)((n + batch - 1) / batch);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  double last = 0.0;
  for (long it = 0; it < steps && !rc; it++) {
    for (int b = 0; b < batch; b++) hidx[b] = (unsigned)(smix(st) % n);
    if (hipMemcpyAsync(g->idx, hidx.data(), 4L * batch, hipMemcpyHostToDevice, g->s) != hipSuccess) { rc = -3; break; }
    hipLaunchKernelGGL(k_surr_batch, dim3((batch + 255) / 256), dim3(256), 0, g->s, batch, g->idx, data, g->R, g->lsc, g->X,
                       g->Tg, g->Wt);
    surr_fwd_bwd(g, batch);
    /* cosine decay of the step size over the run */
    const double lr_it = lr * 0.5 * (1.0 + cos(M_PI * (double)it / (double)steps));
    g->steps++;
    const double bc1 = 1.0 - pow(b1, (double)g->steps), bc2 = 1.0 - pow(b2, (double)g->steps);
    const float lr_t = (float)(lr_it * sqrt(bc2) / bc1);
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((SURR_NPARAM + 255) / 256)), dim3(256), 0, g->s, (long)SURR_NPARAM, g->P, g->G,
                       g->Am, g->Av, lr_t, b1, b2, eps);
    if (it + 1 == steps || (it & 255) == 0) {
      /* hipMemcpyAsync of the indices above needs the host buffer until the copy ran */
      last = surr_loss(g, batch);
    }
    if (hipStreamSynchronize(g->s) != hipSuccess) rc = -3; /* the host index buffer is reused next step */
  }
  (void)hipFree(data);
  if (!rc) rc = surr_publish(g);
  if (!rc && loss_out) *loss_out = last;
  g->validated = false;
  if (rc && !*g_err) set_err("surrogate training failed");
  return rc;
}

/* One forward/backward on a host batch (X[B][3] network inputs, targets and weights [B][34]):
 * returns the loss and the FP32 gradients (SURR_NPARAM, layout above) for checking. */
extern "C" int kqed_surrogate_gradients(kqed_surrogate *g, int B, const float *X3, const float *tgt, const float *wgt,
                                        float *grad, double *loss)
{
  if (!g || B < 1 || B > (1 << 16) || !X3 || !tgt || !wgt || !grad) { set_err("kqed_surrogate_gradients: bad argument"); return -1; }
  int rc = surr_ensure_batch(g, B);
  if (rc) return rc;
  std::vector<float> X((size_t)B * 16, 0.f);
  for (int b = 0; b < B; b++)
    for (int k = 0; k < 3; k++) X[(size_t)b * 16 + k] = X3[b * 3 + k];
  if (hipMemcpy(g->X, X.data(), 4L * B * 16, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(g->Tg, tgt, 4L * B * KQED_NSURR, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(g->Wt, wgt, 4L * B * KQED_NSURR, hipMemcpyHostToDevice) != hipSuccess) {
    set_err("surrogate batch upload failed");
    return -3;
  }
  surr_fwd_bwd(g, B);
  const double l = surr_loss(g, B);
  if (l < 0 || hipMemcpy(grad, g->G, 4 * SURR_NPARAM, hipMemcpyDeviceToHost) != hipSuccess) { set_err("surrogate gradient read failed"); return -3; }
  if (loss) *loss = l;
  return 0;
}

extern "C" int kqed_surrogate_params(const kqed_surrogate *g, float *P, float *lsc, unsigned char *mask, float *range)
{
  if (!g) return -1;
  if (P && hipMemcpy(P, g->P, 4 * SURR_NPARAM, hipMemcpyDeviceToHost) != hipSuccess) return -3;
  if (lsc) memcpy(lsc, g->h_lsc.data(), 4 * g->h_lsc.size());
  if (mask) memcpy(mask, g->h_mask.data(), NBB);
  if (range) { range[0] = g->R.lx0; range[1] = g->R.lx1; range[2] = g->R.ly0; range[3] = g->R.ly1; }
  return 0;
}

extern "C" size_t kqed_surrogate_nparam(void) { return SURR_NPARAM; }

/* n pairs through either path, FP64 assembly: surrogate (every packed pair, no mask) or table */
static int surr_eval(kqed_surrogate *g, size_t n, const double *x_dev, const double *y_dev, double *out, bool use,
                     unsigned char *allmask)
{
  kqed_device *d = g->d;
  SurrUse su = { surr_net(g), allmask };
  std::lock_guard<std::mutex> lk(d->mu);
  kqed_tune keep = d->tune;
  d->tune.sort_min = 0; /* the surrogate runs inside the sorted pipeline only */
  if (d->tune.sort_mode == KQED_SORT_NONE) d->tune.sort_mode = KQED_SORT_CELL;
  const int rc = run_pairs(d, d->ws[(void *)g->s], g->s, KQED_L0, 0.0, n, x_dev, 4, y_dev, 4, out, KQED_AOS, nullptr, nullptr,
                           use ? &su : nullptr);
  d->tune = keep;
  return rc;
}

/* Error check: on validation pairs, the surrogate L0 against the table L0 (both FP64 assembly);
 * a bin passes when it holds at least min_count packed-path pairs and its largest normwise error
 * is at most tol.  stats = {pairs, packed pairs, packed pairs in passing bins, max error, median
 * error over packed pairs}. */
extern "C" int kqed_surrogate_validate(kqed_surrogate *g, size_t n, const double *x_dev, const double *y_dev, double tol,
                                       int min_count, double *stats)
{
  if (!g || !x_dev || !y_dev || n == 0) { set_err("kqed_surrogate_validate: bad argument"); return -1; }
  double *a = nullptr, *b = nullptr, *data = nullptr;
  int *bin = nullptr;
  unsigned *emax = nullptr, *ecnt = nullptr;
  float *err = nullptr;
  unsigned char *all = nullptr;
  int rc = 0;
  if (hipMalloc(&a, 3072 * n) != hipSuccess || hipMalloc(&b, 3072 * n) != hipSuccess ||
      hipMalloc(&data, 8 * KQED_SURR_ROW * n) != hipSuccess || hipMalloc(&bin, 4 * n) != hipSuccess ||
      hipMalloc(&emax, 4 * NBB) != hipSuccess || hipMalloc(&ecnt, 4 * NBB) != hipSuccess ||
      hipMalloc(&err, 4 * n) != hipSuccess || hipMalloc(&all, NBB) != hipSuccess) {
    set_err("surrogate validation buffers failed");
    rc = -2;
  }
  std::vector<unsigned> hm(NBB), hc(NBB);
  std::vector<float> he(n);
  std::vector<int> hb(n);
  if (!rc && (hipMemsetAsync(all, 1, NBB, g->s) != hipSuccess || hipMemsetAsync(emax, 0, 4 * NBB, g->s) != hipSuccess ||
              hipMemsetAsync(ecnt, 0, 4 * NBB, g->s) != hipSuccess))
    rc = -3;
  if (!rc) rc = surr_eval(g, n, x_dev, y_dev, a, true, all);
  if (!rc) rc = surr_eval(g, n, x_dev, y_dev, b, false, nullptr);
  if (!rc) rc = surr_extract(g, n, x_dev, y_dev, data);
  if (!rc) {
    const kqed_dev_tables &T = g->d->dt;
    hipLaunchKernelGGL(k_surr_rowbin, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, g->s, T, n, data, bin);
    hipLaunchKernelGGL(k_surr_binerr, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, g->s, n, a, b, bin, emax, ecnt, err);
    if (hipMemcpyAsync(hm.data(), emax, 4 * NBB, hipMemcpyDeviceToHost, g->s) != hipSuccess ||
        hipMemcpyAsync(hc.data(), ecnt, 4 * NBB, hipMemcpyDeviceToHost, g->s) != hipSuccess ||
        hipMemcpyAsync(he.data(), err, 4 * n, hipMemcpyDeviceToHost, g->s) != hipSuccess ||
        hipMemcpyAsync(hb.data(), bin, 4 * n, hipMemcpyDeviceToHost, g->s) != hipSuccess ||
        hipStreamSynchronize(g->s) != hipSuccess)
      rc = -3;
  }
  (void)hipFree(a); (void)hipFree(b); (void)hipFree(data); (void)hipFree(bin); (void)hipFree(emax); (void)hipFree(ecnt);
  (void)hipFree(err); (void)hipFree(all);
  if (rc) { if (!*g_err) set_err("surrogate validation failed"); return rc; }
  size_t packed = 0, pass = 0;
  std::vector<float> pe;
  for (int c = 0; c < NBB; c++) {
    float m;
    memcpy(&m, &hm[c], 4);
    g->h_mask[c] = hc[c] >= (unsigned)(min_count < 1 ? 1 : min_count) && m <= tol;
  }
  for (size_t p = 0; p < n; p++) {
    if (hb[p] < 0) continue;
    packed++;
    pe.push_back(he[p]);
    if (g->h_mask[hb[p]]) pass++;
  }
  std::sort(pe.begin(), pe.end());
  if (hipMemcpy(g->mask, g->h_mask.data(), NBB, hipMemcpyHostToDevice) != hipSuccess) { set_err("mask upload failed"); return -3; }
  g->validated = true;
  if (stats) {
    stats[0] = (double)n;
    stats[1] = (double)packed;
    stats[2] = (double)pass;
    stats[3] = pe.empty() ? 0.0 : pe.back();
    stats[4] = pe.empty() ? 0.0 : pe[pe.size() / 2];
  }
  return 0;
}

/* kqed_batch with the surrogate: packed-path pairs in bins that passed the error check take the
 * network, every other pair (Taylor region, special points, failing bins) the table path. */
extern "C" int kqed_batch_surrogate(const struct QED_kernel_temps *t, const kqed_surrogate *g, int which, double M,
                                    size_t n, const double *x_dev, const double *y_dev, void *out_dev, int layout,
                                    void *stream)
{
  if (!t || !t->dev || !g || g->d != t->dev) { set_err("kqed_batch_surrogate: bad argument"); return -1; }
  if (which < 0 || which > 4 || !layout_ok(layout)) { set_err("kqed_batch_surrogate: bad argument"); return -1; }
  if (!g->validated) { set_err("kqed_batch_surrogate: surrogate not validated (kqed_surrogate_validate)"); return -1; }
  if (n == 0) return 0;
  if (n > 0xFFFFFFFFull) { set_err("kqed_batch_surrogate: at most 2^32-1 pairs per call"); return -1; }
  kqed_device *d = t->dev;
  SurrUse su = { surr_net(g), g->mask };
  /* no bin passed its error check: every pair takes the tables, so the plain pipeline gives the
   * same result */
  bool any = false;
  for (unsigned char m : g->h_mask) any = any || m;
  std::lock_guard<std::mutex> lk(d->mu);
  if (!any)
    return run_pairs(d, d->ws[stream], (hipStream_t)stream, which, M, n, x_dev, 4, y_dev, 4, out_dev, layout, nullptr,
                     nullptr);
  kqed_tune keep = d->tune;
  d->tune.sort_min = 0;
  if (d->tune.sort_mode == KQED_SORT_NONE) d->tune.sort_mode = KQED_SORT_CELL;
  const int rc = run_pairs(d, d->ws[stream], (hipStream_t)stream, which, M, n, x_dev, 4, y_dev, 4, out_dev, layout, nullptr,
                           nullptr, &su);
  d->tune = keep;
  return rc;
}

/* L0 with the packed-path weight functions supplied by the caller (w_dev[n][34]): the pairs
 * interpolated on the packed path (|x| in [XX[0], XX[last]), post-swap |y| below YY[last]) take
 * w_dev's values instead of the tables and run the chain-rule algebra and assembly unchanged; every
 * other pair is evaluated from the tables.  No counterpart in the reference: it checks the algebra
 * on given inputs, the step the surrogate feeds. */
extern "C" int kqed_batch_inject(const struct QED_kernel_temps *t, size_t n, const double *x_dev, const double *y_dev,
                                 const double *w_dev, void *out_dev, int layout, void *stream)
{
  if (!t || !t->dev || !w_dev || !layout_ok(layout)) { set_err("kqed_batch_inject: bad argument"); return -1; }
  if (n == 0) return 0;
  if (n > 0xFFFFFFFFull) { set_err("kqed_batch_inject: at most 2^32-1 pairs per call"); return -1; }
  kqed_device *d = t->dev;
  std::lock_guard<std::mutex> lk(d->mu);
  if (!d->allmask) {
    std::vector<unsigned char> ones(NBB, 1);
    if (hipMalloc(&d->allmask, NBB) != hipSuccess) { d->allmask = nullptr; set_err("device allocation failed"); return -2; }
    if (hipMemcpy(d->allmask, ones.data(), NBB, hipMemcpyHostToDevice) != hipSuccess) { set_err("mask upload failed"); return -3; }
  }
  SurrUse su = { SurrNet(), d->allmask, w_dev };
  kqed_tune keep = d->tune;
  d->tune.sort_min = 0; /* the substitution runs inside the sorted pipeline only */
  if (d->tune.sort_mode == KQED_SORT_NONE) d->tune.sort_mode = KQED_SORT_CELL;
  const int rc = run_pairs(d, d->ws[stream], (hipStream_t)stream, KQED_L0, 0.0, n, x_dev, 4, y_dev, 4, out_dev, layout,
                           nullptr, nullptr, &su);
  d->tune = keep;
  return rc;
}

/* Save / load: magic, parameter count, parameters, scales, mask, input ranges, validated flag. */
extern "C" int kqed_surrogate_save(const kqed_surrogate *g, const char *path)
{
  if (!g || !path) return -1;
  std::vector<float> P(SURR_NPARAM);
  if (hipMemcpy(P.data(), g->P, 4 * SURR_NPARAM, hipMemcpyDeviceToHost) != hipSuccess) { set_err("surrogate read failed"); return -3; }
  FILE *f = fopen(path, "wb");
  if (!f) { set_err("cannot write %s", path); return -1; }
  const unsigned magic = 0x4B514553u, np = (unsigned)SURR_NPARAM, val = g->validated;
  bool ok = fwrite(&magic, 4, 1, f) == 1 && fwrite(&np, 4, 1, f) == 1 && fwrite(P.data(), 4, P.size(), f) == P.size() &&
            fwrite(g->h_lsc.data(), 4, g->h_lsc.size(), f) == g->h_lsc.size() &&
            fwrite(g->h_mask.data(), 1, g->h_mask.size(), f) == g->h_mask.size() && fwrite(&g->R, sizeof g->R, 1, f) == 1 &&
            fwrite(&val, 4, 1, f) == 1;
  ok = fclose(f) == 0 && ok;
  if (!ok) { set_err("write error on %s", path); return -1; }
  return 0;
}

extern "C" int kqed_surrogate_load(kqed_surrogate *g, const char *path)
{
  if (!g || !path) return -1;
  FILE *f = fopen(path, "rb");
  if (!f) { set_err("cannot read %s", path); return -1; }
  unsigned magic = 0, np = 0, val = 0;
  std::vector<float> P(SURR_NPARAM);
  SurrRange R;
  bool ok = fread(&magic, 4, 1, f) == 1 && fread(&np, 4, 1, f) == 1 && magic == 0x4B514553u && np == SURR_NPARAM &&
            fread(P.data(), 4, P.size(), f) == P.size() && fread(g->h_lsc.data(), 4, g->h_lsc.size(), f) == g->h_lsc.size() &&
            fread(g->h_mask.data(), 1, g->h_mask.size(), f) == g->h_mask.size() && fread(&R, sizeof R, 1, f) == 1 &&
            fread(&val, 4, 1, f) == 1;
  fclose(f);
  if (!ok) { set_err("%s is not a surrogate file of this library", path); return -1; }
  g->R = R;
  g->validated = val != 0;
  if (hipMemcpy(g->P, P.data(), 4 * SURR_NPARAM, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(g->lsc, g->h_lsc.data(), 4 * g->h_lsc.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(g->mask, g->h_mask.data(), g->h_mask.size(), hipMemcpyHostToDevice) != hipSuccess || surr_publish(g)) {
    set_err("surrogate upload failed");
    return -3;
  }
  return 0;
}
