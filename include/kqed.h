/*
 * kqed.h -- C ABI of the MI355X implementation of KQED's position-space QED kernel
 * L_[rho,sigma];mu,nu,lambda(x,y).
 *
 * The first block keeps the reference library's public API (reference src/HEADERS/KQED.h,
 * init.h, kernels.h) so that code written against it recompiles unchanged.  The second block adds
 * batched, device-pointer entry points.  All pointers are plain host or device pointers; no HIP
 * or torch type appears in a signature (streams are passed as void*, a hipStream_t).
 *
 * Output layout everywhere: double kerv[6][4][4][4], flat index 64*rhosig + 16*mu + 4*nu + lambda,
 * rhosig enumerating (0,1),(0,2),(0,3),(1,2),(1,3),(2,3) (reference pi_pert.c:196-199).
 */
#ifndef KQED_AMD_H
#define KQED_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- reference-compatible types */
/* Replaces struct Grid_coeffs (reference KQED.h:69-87).  Host mirrors of the grids are kept for
 * callers that read them; the tables themselves live in device memory behind `dev`. */
struct Grid_coeffs {
  int Nffa;
  double *XX;
  double *YY;
  double xstp;
  double ystp;
  int nstpx;
  int nstpy;
  int NtayY;
  int NY_tay;
  int NtayX;
  int NX_tay;
  int *nfx;
};

/* Replaces struct QED_kernel_temps (reference KQED.h:90-93), passed by value as in the reference. */
struct QED_kernel_temps {
  struct Grid_coeffs Grid;
  int *G8;                     /* 65536 eight-gamma traces, as the reference's init_g8 */
  struct kqed_device *dev;     /* opaque: device-resident tables and assembly maps */
};

/* ---------------------------------------------------------------- reference-compatible API */
/* Loads FFxy_single_cksum.bin, taylorx_cksum.bin and taylory_cksum.bin from $KQED_TABLE_DIR
 * (else the compile-time default directory), verifies their checksums, builds the device tables.
 * Returns 0 on success, 1 on error (message on stderr).  Replaces reference init.c:142-190. */
int initialise(struct QED_kernel_temps *t);
/* Frees host and device memory; safe on a failed or zeroed struct.  Replaces init.c:38-138. */
void free_QED_temps(struct QED_kernel_temps *t);

/* One (x,y) pair each, synchronous, host pointers; overwrite all 384 outputs.  Out-of-grid inputs
 * give an all-zero kernel, as in the reference.  Replace reference kernels.h:4-33 /
 * kernels.c:82-231. */
void QED_kernel_L0(const double xv[4], const double yv[4], const struct QED_kernel_temps t,
                   double kerv[6][4][4][4]);
void QED_kernel_L1(const double xv[4], const double yv[4], const struct QED_kernel_temps t,
                   double kerv[6][4][4][4]);
void QED_kernel_L2(const double xv[4], const double yv[4], const struct QED_kernel_temps t,
                   double kerv[6][4][4][4]);
void QED_kernel_L3(const double xv[4], const double yv[4], const struct QED_kernel_temps t,
                   double kerv[6][4][4][4]);
void QED_Mkernel_L2(const double M, const double xv[4], const double yv[4],
                    const struct QED_kernel_temps t, double kerv[6][4][4][4]);

/* Kernel containers of the combination routines (reference KQED.h:105-131), same layout:
 * 384 doubles per kernel, kernels in declaration order. */
struct Kernel {
  double xy[6][4][4][4];
  double yx[6][4][4][4];
};
struct Kernels {
  double L0[6][4][4][4];
  double L1[6][4][4][4];
  double L2[6][4][4][4];
  double L3[6][4][4][4];
};
struct QED_Kernels {
  struct Kernel L0;
  struct Kernel L1;
  struct Kernel L2;
  struct Kernel L3;
};
struct QED_Con_Kernels {
  struct Kernels A; /* L^i(x,y) */
  struct Kernels B; /* L^i(x,x-y) */
  struct Kernels C; /* L^i(x,y) + L^i_{mu<>nu}(y,x) - L^i(x,x-y) */
};

/* Kernel combinations, one (x,y) pair each, synchronous, host pointers; computed on the GPU
 * through kqed_combine.  Replace reference all_kernels.h:4-33, con_kernel.h:4-27,
 * sub_kernel.h:4-9, SYMXY.h:4-16, SYMXY0.h:4-16 (all_kernels.c, con_kernel.c, sub_kernel.c,
 * SYMXY.c, SYMXY0.c).  On a device error they print the message and leave zeros. */
void compute_all_kernels(const double xv[4], const double yv[4], const struct QED_kernel_temps t,
                         struct QED_Kernels *K);
void compute_all_Mkernels(const double M[4], const double xv[4], const double yv[4],
                          const struct QED_kernel_temps t, struct QED_Kernels *K);
void compute_all_Mkernels_v2(const double M[4], const double xv[4], const double yv[4],
                             const struct QED_kernel_temps t, struct QED_Kernels *K);
void compute_con_kernels(const double xv[4], const double yv[4], const struct QED_kernel_temps t,
                         struct QED_Kernels *K);
void compute_con_kernels_v2(const double xv[4], const double yv[4], const struct QED_kernel_temps t,
                            struct Kernels *K /* [3] */);
void compute_con_kernelsM_L2(const double M[4], const double xv[4], const double yv[4],
                             const struct QED_kernel_temps t, struct Kernels *K /* [3] */);
void compute_sub_kernelsM_L2(const double M[4], const double xv[4], const double yv[4],
                             const struct QED_kernel_temps t, struct Kernels *K /* [3] */);
void compute_all_kernels_SYMXY(const double xv[4], const double yv[4], const struct QED_kernel_temps t,
                               struct QED_Kernels *K);
void compute_all_kernels_SYMXY_v2(const double xv[4], const double yv[4], const struct QED_kernel_temps t,
                                  struct Kernels *K);
void compute_all_kernels_SYMXY0(const double xv[4], const double yv[4], const struct QED_kernel_temps t,
                                struct QED_Kernels *K);
void compute_all_kernels_SYMXY0_v2(const double xv[4], const double yv[4], const struct QED_kernel_temps t,
                                   struct Kernels *K);
/* Host-side helpers on the containers: K = (K + W) / 2, and mu <-> nu of every .yx kernel.
 * Replace all_kernels.c:18-82 and :372-406, con_kernel.c:13-59. */
void average_all_kernels(struct Kernels *K, const struct Kernels W);
void average_all_QED_kernels(struct QED_Kernels *K, const struct QED_Kernels W);
void average_con_kernels(struct Kernels *K /* [3] */, const struct Kernels *W /* [3] */);
void swap_munu_Lyx(struct QED_Kernels *K);

/* Perturbative lepton loop (reference pi_pert.h): the free-lepton four-point function
 * iPihat_{rho sigma; mu nu lambda}(x, y) (x, y in units of the loop lepton's inverse mass; singular
 * where x, y or x-y vanishes) and the analytic light-by-light results it is checked against.
 * ipihatFermLoop_antisym and pihat1 run on the GPU (one pair each, synchronous); the rest is host
 * arithmetic.  Replace pi_pert.c:50-397.  amu_heavlept / amu_lightlept do not print their terms. */
double amu_heavlept(const double mh);
double amu_lightlept(const double ml);
double vac_pol(const double ml);
void bessk(const int n, const double x, double *bv);
int diractrace(const int n, const int *idx);
void init_g8(struct QED_kernel_temps *t);
void ipihatFermLoop_antisym(const double x[4], const double y[4], const struct QED_kernel_temps t,
                            double vpihat[6][4][4][4]);
void pihat1(const double x[4], const double y[4], const struct QED_kernel_temps t,
            double vpihat1[4][4][4][4][4], double vpihatr1[4][4][4][4]);

/* ---------------------------------------------------------------- batched entry points (new) */
enum kqed_kernel { KQED_L0 = 0, KQED_L1 = 1, KQED_L2 = 2, KQED_L3 = 3, KQED_ML2 = 4 };
/* Output layouts.  The FP32 layouts also run the sparse assembly maps in FP32 (interpolation, the
 * Taylor branches and the L1..L3 / M-L2 subtractions stay FP64). */
enum kqed_layout {
  KQED_AOS = 0,     /* double out[n][384] */
  KQED_SOA = 1,     /* double out[384][n] */
  KQED_AOS_F32 = 2, /* float out[n][384] */
  KQED_SOA_F32 = 3  /* float out[384][n] */
};

/* Initialise from an explicit table directory (initialise() with a path argument). */
int kqed_initialise_dir(struct QED_kernel_temps *t, const char *table_dir);

/* n pairs, device pointers: x[n][4], y[n][4] (FP64), out FP64 in `layout`.  Enqueued on `stream`
 * (a hipStream_t; NULL = the null stream); returns 0 or a negative error code without waiting.
 * Replaces the reference GPU port's per-array wrapper cu_arr_QED_kernel_L0
 * (reference src_cuda/amu_forlattice.cu:184-202) and extends it to L1..L3 and M-L2. */
int kqed_batch(const struct QED_kernel_temps *t, int which, double M, size_t n,
               const double *x_dev, const double *y_dev, double *out_dev, int layout, void *stream);

/* Same on host buffers (copies in and out, synchronous); out is double or float per layout. */
int kqed_batch_host(const struct QED_kernel_temps *t, int which, double M, size_t n,
                    const double *x, const double *y, void *out, int layout);

/* Ragged groups: group g pairs y_dev[g] with x_dev[offsets[g] .. offsets[g+1]) (offsets on the
 * host, ngroups+1 entries, offsets[0] = 0); output row i belongs to pair i of the concatenation.
 * Enqueued on `stream`; the offsets are copied before return. */
int kqed_batch_ragged(const struct QED_kernel_temps *t, int which, double M, size_t ngroups,
                      const size_t *offsets, const double *y_dev, const double *x_dev, void *out_dev,
                      int layout, void *stream);

/* The same ragged batch captured once into a HIP graph (own workspace): each kqed_graph_launch
 * replays the whole pipeline as a single graph launch on `stream`.  The device buffers named at
 * creation must stay valid; their contents may change between launches.  Launch parameters are
 * those in force at creation (kqed_set_tune). */
typedef struct kqed_graph kqed_graph;
int kqed_graph_create_ragged(const struct QED_kernel_temps *t, int which, double M, size_t ngroups,
                             const size_t *offsets, const double *y_dev, const double *x_dev,
                             void *out_dev, int layout, kqed_graph **graph);
int kqed_graph_launch(kqed_graph *graph, void *stream);
void kqed_graph_destroy(kqed_graph *graph);

/* Launch parameters of the pipeline (per initialised table set). */
enum kqed_sort_mode { KQED_SORT_CELL = 0, /* by regime and (x node, y node) cell */
                      KQED_SORT_XNODE = 1, /* by regime and x node */
                      KQED_SORT_NONE = 2 };
typedef struct {
  size_t chunk;    /* pairs per pipeline pass (records of one pass stay in device memory) */
  int asm_grid;    /* workgroups of the assembly kernel (each loops over its pairs) */
  int sort_mode;   /* enum kqed_sort_mode */
  size_t sort_min; /* batches below this size are not sorted */
} kqed_tune;
int kqed_get_tune(const struct QED_kernel_temps *t, kqed_tune *tune);
int kqed_set_tune(const struct QED_kernel_temps *t, const kqed_tune *tune);

/* Autotuner: times every candidate (pass length 2^18..2^22, assembly grid 512..4096, sort by
 * cell or x node) on the given batch, `reps` timed runs each, installs the fastest and returns it
 * (and its time per run in ms). */
int kqed_autotune(const struct QED_kernel_temps *t, int which, double M, size_t n, const double *x_dev,
                  const double *y_dev, void *out_dev, int layout, int reps, void *stream, kqed_tune *best,
                  double *best_ms);

/* Contraction with a per-site tensor: for j < ny,
 *   s[j] = sum_{i<nx} sum_{c<384} L(x_i, y_j)_c * T_i[c],   L = L0..L3 or M-L2 per `which`,
 * x_dev[nx][4] FP64 and T_dev[nx][384] FP32 on the device, y[ny][4] and s[ny] FP64 on the host.
 * Accumulation is FP64 with a fixed summation order (deterministic).  Synchronous: returns after s
 * is written, 0 or a negative error code.  The reference has no library routine for this; it is
 * the pattern of its examples (reference amu_forlattice.c:99-107, pihat . kerv summed per point). */
int kqed_contract(const struct QED_kernel_temps *t, int which, double M, size_t nx, const double *x_dev,
                  const float *T_dev, size_t ny, const double *y, double *s, void *stream);

/* kqed_batch that allocates no device memory: no counterpart in the reference; for callers that
 * cannot allocate (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * Its workspace is a fixed scratch array in device memory, one per GPU, held until the process
 * ends (about 43 MB): at most KQED_NOALLOC_MAX_PAIRS pairs per call (larger batches are rejected
 * with an error), evaluated in passes of KQED_NOALLOC_PASS pairs, tables with at most
 * KQED_NOALLOC_MAX_BUCKETS sort buckets (1 + nstpx * nstpy + nstpy).  Outputs equal kqed_batch's.
 * Every call on a GPU shares that GPU's array: calls that overlap in time on one GPU (any stream,
 * any table set) race.  Enqueued on `stream`; 0 or a negative error code. */
#define KQED_NOALLOC_MAX_PAIRS 262144
#define KQED_NOALLOC_PASS 32768
#define KQED_NOALLOC_MAX_BUCKETS (1 + 100 * 4096 + 4096)
int kqed_batch_noalloc(const struct QED_kernel_temps *t, int which, double M, size_t n, const double *x_dev,
                       const double *y_dev, void *out_dev, int layout, void *stream);

/* Bytes of device scratch kqed_batch needs for n pairs (allocated and cached by the library, per
 * stream): kqed_batch_workspace assumes the default pass of 2^21 pairs and a table of at most
 * 100 x 4096 nodes; kqed_batch_workspace_tables uses the table set's own grid and current tune. */
size_t kqed_batch_workspace(size_t n);
size_t kqed_batch_workspace_tables(const struct QED_kernel_temps *t, size_t n);

/* Stage timing of kqed_batch on its own stream (HIP events around every kernel launch).
 * Reads and resets the accumulated milliseconds and launch counts of the stages
 * {0: cell sort, 1: interpolation kernel, 2: assembly kernel} (either pointer may be NULL; arrays
 * of 3), then turns timing on (enable != 0) or off.  Waits for the recorded events. */
int kqed_profile(const struct QED_kernel_temps *t, int enable, double *ms, long *count);

/* Last error text of this thread ("" if none). */
const char *kqed_last_error(void);

/* ---------------------------------------------------------------- learned surrogate (new) */
/* An MLP (3 -> 4 x 128 SiLU -> 34; BF16 weights, FP32 accumulation on the matrix cores) standing in
 * for the table interpolation of the 34 weight functions of packed-path pairs (reference
 * getff-new.c:214-312, simd.c:343-687, as consumed by chnr_dS.c, chnr_dT.c, chnr_dV.c).  It is
 * trained on the GPU from the FP64 interpolator's own output on caller-given pairs, validated per
 * (log|x|, log|y|) bin against the table path, and used only in bins that passed; the Taylor
 * region, special points and failing bins always take the tables.  No reference counterpart. */
typedef struct kqed_surrogate kqed_surrogate;
int kqed_surrogate_create(const struct QED_kernel_temps *t, unsigned long long seed, kqed_surrogate **s);
void kqed_surrogate_destroy(kqed_surrogate *s); /* before free_QED_temps of its tables */
/* Interpolator samples on n device pairs, then `epochs` passes of Adam over batches (cosine step
 * decay from lr); returns the last mean squared error of the normalised outputs in *loss. */
int kqed_surrogate_train(kqed_surrogate *s, size_t n, const double *x_dev, const double *y_dev, int epochs,
                         int batch, double lr, unsigned long long seed, double *loss);
/* Error check on n device pairs: surrogate L0 vs table L0 per pair (normwise); a bin passes with
 * at least min_count packed pairs all within tol.  stats[5] = {pairs, packed pairs, packed pairs in
 * passing bins, max error, median error}. */
int kqed_surrogate_validate(kqed_surrogate *s, size_t n, const double *x_dev, const double *y_dev,
                            double tol, int min_count, double *stats);
/* kqed_batch with the surrogate in the bins that passed (requires a validated surrogate). */
int kqed_batch_surrogate(const struct QED_kernel_temps *t, const kqed_surrogate *s, int which, double M,
                         size_t n, const double *x_dev, const double *y_dev, void *out_dev, int layout,
                         void *stream);
/* L0 of n device pairs with the packed-path weight functions given by the caller instead of the
 * tables: w_dev[n][34] per input pair, in the order g0: dcb dx dy | l2: dx dy val dcb | l4: dx dy
 * val dcb | v1: dx dy val dcb | g2: val dy | dcb dydcb dcbdcb | dxdx dx dxdy dxdcb | g3: dxdx dx dxdy
 * dxdcb | dcb dydcb dcbdcb | dy | swap regime only: g2 dydy, g3 dydy, each a derivative in the
 * post-swap invariants (|x|, y', cb') of the weight function chnr_dS.c:50-95, chnr_dT.c:185-293 and
 * chnr_dV.c:363-479 read from extractff2/extractff.  Pairs not on the packed path (Taylor region,
 * special points, grid edges) ignore their row.  The chain-rule algebra and assembly then run
 * unchanged (the step the surrogate feeds).  No reference counterpart: it exposes that step for
 * checking.  Enqueued on `stream`; 0 or a negative error code. */
int kqed_batch_inject(const struct QED_kernel_temps *t, size_t n, const double *x_dev, const double *y_dev,
                      const double *w_dev, void *out_dev, int layout, void *stream);
/* One forward/backward pass on a host batch (inputs X[B][3], targets and weights [B][34]): loss and
 * the FP32 gradient of every parameter (kqed_surrogate_nparam() floats), for checking. */
int kqed_surrogate_gradients(kqed_surrogate *s, int B, const float *X, const float *target,
                             const float *weight, float *grad, double *loss);
size_t kqed_surrogate_nparam(void);
/* Parameters (FP32), log output scales [34][24][24], bin mask [24*24], input ranges [4]; any
 * pointer may be NULL. */
int kqed_surrogate_params(const kqed_surrogate *s, float *params, float *log_scales, unsigned char *mask,
                          float *range);
int kqed_surrogate_save(const kqed_surrogate *s, const char *path);
int kqed_surrogate_load(kqed_surrogate *s, const char *path);

/* ---------------------------------------------------------------- table utilities (new) */
/* Write a synthetic form-factor table in the reference's file format (see kqed_tables.h). */
int kqed_write_synthetic_ff(const char *path, int nstpx, int nstpy, int nf0, int nf1, double xh,
                            double yh, unsigned long long seed);

/* Batched kernel combinations: n pairs, device pointers x[n][4], y[n][4]; out[n][len] doubles with
 * len = 384 * kqed_comb_kernels(comb), per pair the reference's struct layout (QED_Kernels,
 * Kernels or Kernels[3]).  M: host pointer to 4 doubles for the M variants (ignored otherwise,
 * NULL allowed).  Every L0 the combination needs runs through the batched pipeline; the
 * workspace holds one pass of tune.chunk L0 tensors (3 KiB each).  Enqueued on `stream`. */
enum kqed_comb {
  KQED_COMB_ALL = 0,      /* compute_all_kernels           -> struct QED_Kernels */
  KQED_COMB_ALL_M = 1,    /* compute_all_Mkernels          -> struct QED_Kernels */
  KQED_COMB_ALL_M_V2 = 2, /* compute_all_Mkernels_v2       -> struct QED_Kernels */
  KQED_COMB_CON = 3,      /* compute_con_kernels           -> struct QED_Kernels */
  KQED_COMB_CON_V2 = 4,   /* compute_con_kernels_v2        -> struct Kernels[3] */
  KQED_COMB_CON_M_L2 = 5, /* compute_con_kernelsM_L2       -> struct Kernels[3] */
  KQED_COMB_SUB_M_L2 = 6, /* compute_sub_kernelsM_L2       -> struct Kernels[3] */
  KQED_COMB_SYMXY = 7,    /* compute_all_kernels_SYMXY     -> struct QED_Kernels */
  KQED_COMB_SYMXY_V2 = 8, /* compute_all_kernels_SYMXY_v2  -> struct Kernels */
  KQED_COMB_SYMXY0 = 9,   /* compute_all_kernels_SYMXY0    -> struct QED_Kernels */
  KQED_COMB_SYMXY0_V2 = 10, /* compute_all_kernels_SYMXY0_v2 -> struct Kernels */
  KQED_COMB_N = 11
};
/* kernels of 384 doubles per pair in the output of `comb`: 8 (QED_Kernels), 12 (Kernels[3]) or
 * 4 (Kernels); 0 for an unknown comb */
int kqed_comb_kernels(int comb);
int kqed_combine(const struct QED_kernel_temps *t, int comb, const double *M, size_t n,
                 const double *x_dev, const double *y_dev, double *out_dev, void *stream);
/* Contraction through a combination, for a lattice driver: s[j * nk + q] = sum over the nx sites
 * of sum_c K_q(x_i, y_j)_c T_i[c] for every output kernel q (nk = kqed_comb_kernels(comb)) and y
 * point j; x_dev [nx][4] FP64 and T_dev [nx][384] FP32 on the device, y_host [ny][4] and s_host on
 * the host.  FP64 sums in a fixed order (deterministic).  Synchronous on `stream`. */
int kqed_contract_comb(const struct QED_kernel_temps *t, int comb, const double *M, size_t nx,
                       const double *x_dev, const float *T_dev, size_t ny, const double *y_host,
                       double *s_host, void *stream);
/* Same on host buffers (copies in and out, synchronous). */
int kqed_combine_host(const struct QED_kernel_temps *t, int comb, const double *M, size_t n,
                      const double *x, const double *y, double *out);

/* Table ingestion checks (host only, no GPU).  kqed_tables_info reads the three table files of a
 * directory with the same readers and checksum verification as initialise() and reports their
 * shape; kqed_tables_rewrite writes them back in the reference's format (the writers of reference
 * io.c:292-413, which write fixed names under ./PRECOMP) into dst_dir, for byte-level round trips.
 * Return 0, or the reader's code: 1 missing file, 2 bad magic, 3 bad dimension, 4 checksum
 * mismatch, 5 short read; 6 write error. */
struct kqed_table_info {
  int nstpx, nstpy, nffa;
  int nfmin, nfmax;          /* Chebyshev coefficients per x node (nfx) */
  double x_first, x_last;    /* XX[0], XX[nstpx-1] */
  double y_first, y_last;    /* YY[0], YY[nstpy-1] */
  int x_increasing, y_increasing;
  size_t nonfinite;          /* non-finite form-factor coefficients */
};
int kqed_tables_info(const char *dir, struct kqed_table_info *info);
int kqed_tables_rewrite(const char *src_dir, const char *dst_dir);

/* Lepton loop, batched (device pointers, enqueued on `stream`):
 *   kqed_pihat:       out[n][384] = iPihat(x_i, y_i)  (ipihatFermLoop_antisym per pair)
 *   kqed_lepton_loop: out[n] = sum_c L(x_i, y_i)_c iPihat(Mv x_i, Mv y_i)_c, the integrand of the
 *                     reference's example1 (amu_forlattice.c:66-126) before its normalisation,
 *                     with L = L0..L3 / M-L2 (`which`, `M`) and Mv = lepton mass / muon mass. */
int kqed_pihat(const struct QED_kernel_temps *t, size_t n, const double *x_dev, const double *y_dev,
               double *out_dev, void *stream);
int kqed_pihat_host(const struct QED_kernel_temps *t, size_t n, const double *x, const double *y, double *out);
int kqed_lepton_loop(const struct QED_kernel_temps *t, int which, double M, double Mv, size_t n,
                     const double *x_dev, const double *y_dev, double *out_dev, void *stream);

#ifdef __cplusplus
}
#endif
#endif
