/*
 * kqed_pert.hip -- the perturbative lepton loop iPihat_{rho sigma; mu nu lambda}(x, y) on the GPU
 * (reference pi_pert.c), the four-point function the reference's example1 contracts with L0.
 *
 * ipihatFermLoop_antisym (pi_pert.c:180-220) combines three evaluations of pihat1 (:222-381) at
 * (x, y), (y-x, -x) and (-x, y-x).  Here
 *   k_ipihat   256-thread blocks, each looping over pairs: per pair the three pihat1 evaluations,
 *              each with its propagator factors and gather operands in LDS and thread i =
 *              (sigma, mu, nu, lambda) forming its 4 + 1 outputs from the reference's sum over
 *              (alpha, beta, gamma, delta) regrouped into 16 + 1 + 8 partial sums (each a fixed
 *              number of gathers of the nonzero trace terms, lists built from G8 once), into LDS;
 *              then the (rho, sigma)-antisymmetrised combination of the three, 384 outputs
 *   k_pihat1   the same loop for pihat1 alone (the reference's per-pair pihat1), outputs to HBM
 */
#include "kqed_pert.h"

#include <hip/hip_runtime.h>
#include <math.h>

namespace {

/* Gather lists (built on the host from G8, pi_pert.c:161-177).  Thread t = 64 sigma + 16 mu + 4 nu
 * + lambda of a block forms, for its (sigma, mu, nu, lambda) and with G8, G6, G4 the traces of
 * eight, six and four gammas,
 *   K[gamma delta] = sum_{alpha beta} G8(alpha, mu, beta, nu, gamma, sigma, delta, lambda) xab[alpha beta]
 *                    - G6(mu, nu, gamma, sigma, delta, lambda) b11
 *   A              = sum_{alpha beta} G6(alpha, mu, beta, nu, sigma, lambda) xab[alpha beta] - G4(mu, nu, sigma, lambda) b11
 *   H2[beta]       = sum_alpha G6(alpha, mu, nu, beta, sigma, lambda) x2[alpha] - G6(mu, alpha, nu, beta, sigma, lambda) x3[alpha]
 *   H4[beta]       = sum_alpha G6(alpha, mu, nu, sigma, beta, lambda) x2[alpha] - G6(mu, alpha, nu, sigma, beta, lambda) x3[alpha]
 * and from them
 *   pihat1[rho] = -norm (sum_{gamma delta} fv[gamma delta][rho] K[gamma delta] + g[rho] A
 *                        + sum_beta h[rho][beta] H2[beta] + sum_beta fhat[beta][rho] H4[beta])
 *   pihatr1     = -norm (sum_{gamma delta} ell[gamma delta] K[gamma delta] + p A + sum_beta q[beta] (H2[beta] + H4[beta]))
 * -- the reference's sum over (alpha, beta, gamma, delta) (pi_pert.c:320-377) with its products
 * regrouped (x4 = x2 and x5 = x3 there).  A trace of Euclidean gammas is 0 or +-4, and for one
 * thread at most PT_NK of the 17 terms of a K or of A, and PT_NH of the 8 of an H2 or H4, are nonzero;
 * so every sum is a fixed number of gathers from the evaluation's operand array V in LDS (4 xab,
 * -4 xab, 4 x2, -4 x2, 4 x3, -4 x3, 4 b11, -4 b11, 0).  A thread's list holds the byte offsets into
 * V of its terms, padded with the zero operand, as PT_ENT one-byte slot indices in the order K[0..15]
 * (PT_NK each), A (PT_NK), H2[0..3], H4[0..3] (PT_NH each); the table is laid out
 * [PT_ENT / 16][256 threads][16 entries] so that a wavefront reads its lists in 16-byte rows. */
constexpr int PT_NK = 5, PT_NH = 2, PT_ENT = 112;
enum { V_XAB = 0, V_NXAB = 16, V_X2 = 32, V_NX2 = 36, V_X3 = 40, V_NX3 = 44, V_B11 = 48, V_NB11 = 49, V_ZERO = 50,
       V_LEN = 52 };
static_assert(17 * PT_NK + 8 * PT_NH <= PT_ENT && PT_ENT % 16 == 0, "gather list length");

/* constants of one pihat1 evaluation in LDS */
struct PiConst {
  double V[V_LEN];  /* gather operands; the list offsets count from here */
  double fv[64];    /* fv[gamma][delta][rho] */
  double ell[16];   /* ell[gamma][delta] */
  double h[16];     /* h[rho][delta] */
  double fhat[16];  /* fhat[delta][rho] */
  double g[4], q[4];
  double p;
};

/* v[i] for a thread-dependent i by selects (an indexed register array would go to scratch) */
__device__ inline double comp(const double (&v)[4], int i) { return i == 0 ? v[0] : i == 1 ? v[1] : i == 2 ? v[2] : v[3]; }

__device__ inline double norm4(const double *v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + v[3] * v[3]); }

/* K_0, K_1, K_2 at x (pi_pert.c:92-133): Abramowitz & Stegun 9.8.5-9.8.8 polynomial forms of
 * K0 and K1 and the recurrence K_2 = K_0 + (2/x) K_1 */
__device__ void bessk2(double x, double *bv)
{
  double bkm, bk;
  const double tox = 2.0 / x;
  if (x <= 2.0) {
    double y = x / 3.75;
    y *= y;
    const double i0 = 1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 + y * (0.2659732 + y * (0.360768e-1 + y * 0.45813e-2)))));
    const double i1 = x * (0.5 + y * (0.87890594 + y * (0.51498869 + y * (0.15084934 + y * (0.2658733e-1 + y * (0.301532e-2 + y * 0.32411e-3))))));
    y = 0.25 * x * x;
    const double l = log(x / 2.0);
    bkm = (-l * i0) + (-0.57721566 + y * (0.42278420 + y * (0.23069756 + y * (0.3488590e-1 + y * (0.262698e-2 + y * (0.10750e-3 + y * 0.74e-5))))));
    bk = (l * i1) + (1.0 / x) * (1.0 + y * (0.15443144 + y * (-0.67278579 + y * (-0.18156897 + y * (-0.1919402e-1 + y * (-0.110404e-2 + y * (-0.4686e-4)))))));
  } else {
    const double e = (exp(-x) / sqrt(x));
    bkm = e * (1.25331414 + tox * (-0.7832358e-1 + tox * (0.2189568e-1 + tox * (-0.1062446e-1 + tox * (0.587872e-2 + tox * (-0.251540e-2 + tox * 0.53208e-3))))));
    bk = e * (1.25331414 + tox * (0.23498619 + tox * (-0.3655620e-1 + tox * (0.1504268e-1 + tox * (-0.780353e-2 + tox * (0.325614e-2 + tox * (-0.68245e-3)))))));
  }
  bv[0] = bkm;
  bv[1] = bk;
  bv[2] = bkm + 1 * tox * bk;
}

/* the three argument pairs of ipihatFermLoop_antisym: (x, y), (y-x, -x), (-x, y-x) */
__device__ void pert_args(int k, const double *x, const double *y, double *a, double *b)
{
  for (int i = 0; i < 4; i++) {
    const double mx = -x[i], ymx = y[i] - x[i];
    a[i] = k == 0 ? x[i] : k == 1 ? ymx : mx;
    b[i] = k == 0 ? y[i] : k == 1 ? mx : ymx;
  }
}

/* Bessel functions of the nev evaluations of the ng pairs p0 .. p0 + ng - 1 (pi_pert.c:239-242):
 * thread 9 j + 3 ev + w takes K0, K1 / |v|, K2 / |v|^2 at v = b, a, a - b of evaluation ev's
 * arguments (a, b) of pair p0 + j, all in parallel, into bes[j][ev][w] */
__device__ __forceinline__ void pihat1_bessel(size_t p0, int ng, int nev, const double *__restrict__ X,
                                              const double *__restrict__ Y, double (*bes)[3][3][3])
{
  const int t = threadIdx.x;
  if (t < 9 * ng && t % 9 < 3 * nev) {
    const int j = t / 9, ev = (t % 9) / 3, w = t % 3;
    const size_t pr = p0 + j;
    double px[4], py[4], a[4], b[4], v[4];
    for (int i = 0; i < 4; i++) { px[i] = X[4 * pr + i]; py[i] = Y[4 * pr + i]; }
    pert_args(ev, px, py, a, b);
    for (int i = 0; i < 4; i++) v[i] = w == 0 ? b[i] : w == 1 ? a[i] : a[i] - b[i];
    const double vn = norm4(v);
    double k[3];
    bessk2(vn, k);
    bes[j][ev][w][0] = k[0];
    bes[j][ev][w][1] = k[1] / vn;
    bes[j][ev][w][2] = k[2] / (vn * vn);
  }
}

/* Evaluation ev of pihat1 (0: (x, y); 1, 2: the other arguments of ipihatFermLoop_antisym) at
 * pair pr by the block into o[1280] = v1[256 rho + 64 mu + 16 nu + 4 lambda + sigma], then
 * r1[64 mu + 16 nu + 4 lambda + sigma], from its Bessel values bes; e is the thread's gather list.
 * The tensors go to C before the first barrier and are read after it; the second barrier lets
 * the next evaluation refill C. */
__device__ __forceinline__ void pihat1_item(size_t pr, int ev, const double *__restrict__ X, const double *__restrict__ Y,
                                            const unsigned (&e)[PT_ENT / 4], const double (*bes)[3], PiConst &C,
                                            double *__restrict__ o)
{
  const int t = threadIdx.x;
  double x[4], y[4], xmy[4];
  {
    double px[4], py[4];
    for (int i = 0; i < 4; i++) { px[i] = X[4 * pr + i]; py[i] = Y[4 * pr + i]; }
    pert_args(ev, px, py, x, y);
  }
  for (int i = 0; i < 4; i++) xmy[i] = x[i] - y[i];
  /* the evaluation's tensors (pi_pert.c:244-312), one entry per thread */
  const double pipi = M_PI * M_PI, twopipi = 2 * pipi;
  const double by0 = bes[0][0], by1 = bes[0][1], by2 = bes[0][2];
  if (t < 64) {
    const int c = t >> 4, d = (t >> 2) & 3, r = t & 3;
    const double yc = comp(y, c), yd = comp(y, d), yr = comp(y, r);
    C.fv[t] = pipi * (by2 * yc * yd * yr + ((r == d) * yc - (c == r) * yd - (c == d) * yr) * by1);
  } else if (t < 128) {
    const int u = (t - 64) >> 4, r = (t >> 2) & 3, d = t & 3;
    const double p = twopipi * by0, yr = comp(y, r), yd = comp(y, d);
    if (u == 0) { /* ell[r][d] = e[r] y[d] - delta s */
      const double ee = twopipi * yr * by2, s = twopipi * by1;
      C.ell[4 * r + d] = r == d ? ee * yd - s : ee * yd;
    } else if (u == 1) { /* h[r][d] = f[r] y[d] - delta p/2 */
      const double f = pipi * yr * by1;
      C.h[4 * r + d] = r == d ? f * yd - 0.5 * p : f * yd;
    } else if (u == 2) { /* fhat[d][r] = h[r][d] (+ p on the diagonal) */
      const double f = pipi * yr * by1;
      const double hv = r == d ? f * yd - 0.5 * p : f * yd;
      C.fhat[4 * d + r] = r == d ? hv + p : hv;
    } else { /* xab[r][d] and its gather operands */
      const double v = comp(x, r) * comp(xmy, d) * (bes[1][2] * bes[2][2]);
      C.V[V_XAB + 4 * r + d] = 4.0 * v;
      C.V[V_NXAB + 4 * r + d] = -4.0 * v;
    }
  } else if (t < 132) {
    const int a = t & 3;
    const double v = comp(x, a) * (bes[1][2] * bes[2][1]);
    C.V[V_X2 + a] = 4.0 * v;
    C.V[V_NX2 + a] = -4.0 * v;
  } else if (t < 136) {
    const int a = t & 3;
    const double v = comp(xmy, a) * (bes[1][1] * bes[2][2]);
    C.V[V_X3 + a] = 4.0 * v;
    C.V[V_NX3 + a] = -4.0 * v;
  } else if (t < 140) {
    const int r = t & 3;
    const double yr = comp(y, r);
    C.g[r] = pipi * yr * by0;
    C.q[r] = 2 * (pipi * yr * by1);
  } else if (t == 140) {
    C.p = twopipi * by0;
    const double b11 = bes[1][1] * bes[2][1];
    C.V[V_B11] = 4.0 * b11;
    C.V[V_NB11] = -4.0 * b11;
    C.V[V_ZERO] = 0.0;
    C.V[V_ZERO + 1] = 0.0;
  }
  __syncthreads();
  int j = 0;
  auto gat = [&]() -> double {
    const unsigned slot = (e[j >> 2] >> (8 * (j & 3))) & 0xFFu;
    j++;
    return C.V[slot];
  };
  double K[16];
#pragma unroll
  for (int cd = 0; cd < 16; cd++) {
    double s = gat();
#pragma unroll
    for (int k = 1; k < PT_NK; k++) s += gat();
    K[cd] = s;
  }
  double A = gat();
#pragma unroll
  for (int k = 1; k < PT_NK; k++) A += gat();
  double H2[4], H4[4];
#pragma unroll
  for (int b = 0; b < 4; b++) {
    H2[b] = gat();
#pragma unroll
    for (int k = 1; k < PT_NH; k++) H2[b] += gat();
  }
#pragma unroll
  for (int b = 0; b < 4; b++) {
    H4[b] = gat();
#pragma unroll
    for (int k = 1; k < PT_NH; k++) H4[b] += gat();
  }
  double t1[4] = { 0.0, 0.0, 0.0, 0.0 }, tr1 = 0.0;
#pragma unroll
  for (int cd = 0; cd < 16; cd++) {
#pragma unroll
    for (int r = 0; r < 4; r++) t1[r] = fma(C.fv[4 * cd + r], K[cd], t1[r]);
    tr1 = fma(C.ell[cd], K[cd], tr1);
  }
#pragma unroll
  for (int r = 0; r < 4; r++) {
    t1[r] = fma(C.g[r], A, t1[r]);
#pragma unroll
    for (int b = 0; b < 4; b++) t1[r] = fma(C.h[4 * r + b], H2[b], t1[r]);
#pragma unroll
    for (int b = 0; b < 4; b++) t1[r] = fma(C.fhat[4 * b + r], H4[b], t1[r]);
  }
  tr1 = fma(C.p, A, tr1);
#pragma unroll
  for (int b = 0; b < 4; b++) tr1 = fma(C.q[b], H2[b] + H4[b], tr1);
  const double norm = 8.2336243479292e-07; /* 2 / (2 pi)^8, pi_pert.c:229 */
  const int sg = t >> 6, lane = t & 63;
  const int mu = lane >> 4, nu = (lane >> 2) & 3, ld = lane & 3;
  const int ix = 64 * mu + 16 * nu + 4 * ld + sg;
#pragma unroll
  for (int r = 0; r < 4; r++) o[256 * r + ix] = -norm * t1[r];
  o[1024 + ix] = -norm * tr1;
  __syncthreads();
}

/* the gather lists of thread t, into registers */
__device__ __forceinline__ void load_lists(const unsigned *__restrict__ tab, unsigned (&e)[PT_ENT / 4])
{
  const uint4 *tp = (const uint4 *)tab + threadIdx.x;
#pragma unroll
  for (int r = 0; r < PT_ENT / 16; r++) {
    const uint4 v = tp[256 * r];
    e[4 * r] = v.x; e[4 * r + 1] = v.y; e[4 * r + 2] = v.z; e[4 * r + 3] = v.w;
  }
}

/* pihat1 at pairs blockIdx.x, + gridDim.x, ...: out[pair][1280] */
__global__ void __launch_bounds__(256) k_pihat1(size_t n, const double *__restrict__ X, const double *__restrict__ Y,
                                                const unsigned *__restrict__ tab, double *__restrict__ out)
{
  __shared__ PiConst C;
  __shared__ double bes[1][3][3][3];
  unsigned e[PT_ENT / 4];
  load_lists(tab, e);
  for (size_t pr = blockIdx.x; pr < n; pr += gridDim.x) {
    pihat1_bessel(pr, 1, 1, X, Y, bes);
    __syncthreads();
    pihat1_item(pr, 0, X, Y, e, bes[0][0], C, out + pr * 1280);
  }
}

/* ipihatFermLoop_antisym (pi_pert.c:180-220) over groups of PT_GROUP consecutive pairs, group
 * blockIdx.x, + gridDim.x, ...: the Bessel values of a group's pairs at once (one latency per
 * group), then per pair the three pihat1 evaluations into LDS and the (rho, sigma)-antisymmetrised
 * combination (pi_pert.c:196-218) from there into out[pair][384], rows (rho, sigma) = 01 02 03 12
 * 13 23.  A group's Bessel values are written after every thread has passed the last evaluation
 * of the previous group (whose tensor phase read them); P is read by the combination before the
 * next evaluation's first barrier. */
constexpr int PT_GROUP = 16;
__global__ void __launch_bounds__(256, 3) k_ipihat(size_t n, const double *__restrict__ X, const double *__restrict__ Y,
                                                   const unsigned *__restrict__ tab, double *__restrict__ out)
{
  __shared__ PiConst C;
  __shared__ double bes[PT_GROUP][3][3][3];
  __shared__ double P[3 * 1280];
  const int t = threadIdx.x;
  unsigned e[PT_ENT / 4];
  load_lists(tab, e);
  for (size_t p0 = (size_t)blockIdx.x * PT_GROUP; p0 < n; p0 += (size_t)gridDim.x * PT_GROUP) {
    const int ng = n - p0 < (size_t)PT_GROUP ? (int)(n - p0) : PT_GROUP;
    pihat1_bessel(p0, ng, 3, X, Y, bes);
    __syncthreads();
    for (int j = 0; j < ng; j++) {
      const size_t pr = p0 + j;
#pragma unroll 1
      for (int ev = 0; ev < 3; ev++) pihat1_item(pr, ev, X, Y, e, bes[j][ev], C, P + 1280 * ev);
      double x[4];
      for (int i = 0; i < 4; i++) x[i] = X[4 * pr + i];
      const double *v1 = P, *v2 = P + 1280, *v3 = P + 2560;
      const double *r2 = v2 + 1024, *r3 = v3 + 1024;
#define V5(A, a, b, c, d, e) A[256 * (a) + 64 * (b) + 16 * (c) + 4 * (d) + (e)]
#define R4(A, b, c, d, e) A[64 * (b) + 16 * (c) + 4 * (d) + (e)]
      for (int oi = t; oi < 384; oi += 256) {
        const int rs = oi >> 6, mu = (oi >> 4) & 3, nu = (oi >> 2) & 3, ld = oi & 3;
        const int rho = rs < 3 ? 0 : rs < 5 ? 1 : 2, sig = rs < 3 ? rs + 1 : rs < 5 ? rs - 1 : 3;
        const double xr = comp(x, rho), xs = comp(x, sig);
        out[pr * 384 + oi] =
            0.5 * (+V5(v1, rho, mu, nu, ld, sig) + V5(v2, rho, nu, ld, mu, sig) + xr * R4(r2, nu, ld, mu, sig) +
                   V5(v3, rho, ld, nu, mu, sig) + xr * R4(r3, ld, nu, mu, sig) -
                   (+V5(v1, sig, mu, nu, ld, rho) + V5(v2, sig, nu, ld, mu, rho) + xs * R4(r2, nu, ld, mu, rho) +
                    V5(v3, sig, ld, nu, mu, rho) + xs * R4(r3, ld, nu, mu, rho)));
      }
#undef V5
#undef R4
    }
  }
}

/* scaled copies of the pair arrays: Xs = s x, Ys = s y */
__global__ void k_scale_pairs(size_t n, double s, const double *__restrict__ x, const double *__restrict__ y,
                              double *__restrict__ xs, double *__restrict__ ys)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 4 * n) return;
  xs[i] = s * x[i];
  ys[i] = s * y[i];
}

} // namespace

size_t kqed_pert_table_bytes(void) { return (size_t)256 * PT_ENT; }

int kqed_pert_build_table(const int *G8, unsigned *tab)
{
  /* G8 index of (i0..i7) = sum_k i_k 4^(7-k); six- and four-gamma traces are G8 with leading 0, 0 */
  auto g8 = [G8](int a, int b, int c, int d, int e, int f, int g, int h) {
    return G8[h + 4 * (g + 4 * (f + 4 * (e + 4 * (d + 4 * (c + 4 * (b + 4 * a))))))];
  };
  for (int t = 0; t < 256; t++) {
    const int sg = t >> 6, mu = (t >> 4) & 3, nu = (t >> 2) & 3, ld = t & 3;
    unsigned char L[PT_ENT];
    int n = 0, ng = 0, grp[17];
    bool ok = true;
    /* a term trace * operand: the operand's slot with the trace's sign (traces are 0 or +-4) */
    auto put = [&](int g, int pos, int neg) {
      if (g == 4) grp[ng++] = pos;
      else if (g == -4) grp[ng++] = neg;
      else if (g != 0) ok = false;
    };
    auto flush = [&](int m) {
      if (ng > m) ok = false;
      for (int k = 0; k < m; k++) L[n++] = (unsigned char)(k < ng ? grp[k] : V_ZERO);
      ng = 0;
    };
    for (int c = 0; c < 4; c++)
      for (int d = 0; d < 4; d++) {
        for (int a = 0; a < 4; a++)
          for (int b = 0; b < 4; b++) put(g8(a, mu, b, nu, c, sg, d, ld), V_XAB + 4 * a + b, V_NXAB + 4 * a + b);
        put(-g8(0, 0, mu, nu, c, sg, d, ld), V_B11, V_NB11);
        flush(PT_NK);
      }
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) put(g8(0, 0, a, mu, b, nu, sg, ld), V_XAB + 4 * a + b, V_NXAB + 4 * a + b);
    put(-g8(0, 0, 0, 0, mu, nu, sg, ld), V_B11, V_NB11);
    flush(PT_NK);
    for (int b = 0; b < 4; b++) {
      for (int a = 0; a < 4; a++) {
        put(g8(0, 0, a, mu, nu, b, sg, ld), V_X2 + a, V_NX2 + a);
        put(-g8(0, 0, mu, a, nu, b, sg, ld), V_X3 + a, V_NX3 + a);
      }
      flush(PT_NH);
    }
    for (int b = 0; b < 4; b++) {
      for (int a = 0; a < 4; a++) {
        put(g8(0, 0, a, mu, nu, sg, b, ld), V_X2 + a, V_NX2 + a);
        put(-g8(0, 0, mu, a, nu, sg, b, ld), V_X3 + a, V_NX3 + a);
      }
      flush(PT_NH);
    }
    if (!ok) return -1;
    while (n < PT_ENT) L[n++] = (unsigned char)V_ZERO;
    for (int k = 0; k < PT_ENT; k += 4)
      tab[((size_t)(k / 16) * 256 + t) * 4 + (k % 16) / 4] =
          (unsigned)L[k] | (unsigned)L[k + 1] << 8 | (unsigned)L[k + 2] << 16 | (unsigned)L[k + 3] << 24;
  }
  return 0;
}

/* blocks of the pair loops: at most eight per CU of the 256 */
static unsigned pert_grid(size_t n) { return (unsigned)(n < 2048 ? n : 2048); }
static unsigned pert_groups(size_t n)
{
  const size_t g = (n + PT_GROUP - 1) / PT_GROUP;
  return (unsigned)(g < 2048 ? g : 2048);
}

void kqed_pert_launch_pihat1(size_t n, const double *x, const double *y, const unsigned *tab, double *out, hipStream_t s)
{
  hipLaunchKernelGGL(k_pihat1, dim3(pert_grid(n)), dim3(256), 0, s, n, x, y, tab, out);
}

void kqed_pert_launch_ipihat(size_t n, const double *x, const double *y, const unsigned *tab, double *out, hipStream_t s)
{
  hipLaunchKernelGGL(k_ipihat, dim3(pert_groups(n)), dim3(256), 0, s, n, x, y, tab, out);
}

void kqed_pert_launch_scale(size_t n, double sc, const double *x, const double *y, double *xs, double *ys, hipStream_t s)
{
  hipLaunchKernelGGL(k_scale_pairs, dim3((unsigned)((4 * n + 255) / 256)), dim3(256), 0, s, n, sc, x, y, xs, ys);
}
