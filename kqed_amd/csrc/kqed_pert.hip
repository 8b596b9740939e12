/*
 * kqed_pert.hip -- the perturbative lepton loop iPihat_{rho sigma; mu nu lambda}(x, y) on the GPU
 * (reference pi_pert.c), the four-point function the reference's example1 contracts with L0.
 *
 * ipihatFermLoop_antisym (pi_pert.c:180-220) combines three evaluations of pihat1 (:222-381) at
 * (x, y), (y-x, -x) and (-x, y-x).  Here
 *   k_pihat1         one 256-thread block per (pair, evaluation): the per-evaluation tensors
 *                    (propagator factors, 84 doubles) in LDS, thread i = (sigma, mu, nu, lambda)
 *                    sums its 4 + 1 outputs over (alpha, beta, gamma, delta) in the reference's
 *                    order; the trace factors come from int8 tables laid out so that the 64 lanes
 *                    of a wavefront (fixed sigma) read 64 consecutive bytes per (alpha..delta)
 *   k_pihat_antisym  one 384-thread block per pair: the (rho, sigma)-antisymmetrised combination
 * The traces of eight Euclidean gammas are 4 * {-1, 0, 1}, so int8 holds them exactly.
 */
#include "kqed_pert.h"

#include <hip/hip_runtime.h>
#include <math.h>

namespace {

/* Trace tables (built on the host from G8, pi_pert.c:161-177):
 *   main [alpha][beta][gamma][delta][sigma][lane]  = G8(alpha, mu, beta, nu, gamma, sigma, delta, lambda)
 *   six  [alpha][beta][sigma][k][lane], k = 0..5:  the six-gamma traces of pi_pert.c:345-354
 *   four [sigma][lane]                          = G4(mu, nu, sigma, lambda)
 * with lane = 16 mu + 4 nu + lambda. */
constexpr int PT_MAIN = 65536, PT_SIX = 4 * 4 * 4 * 6 * 64, PT_FOUR = 256;

/* constants of one pihat1 evaluation in LDS */
struct PiConst {
  double fv[64];   /* fv[gamma][delta][rho] */
  double ell[16];  /* ell[gamma][delta] */
  double h[16];    /* h[rho][delta] */
  double fhat[16]; /* fhat[delta][rho] */
  double xab[16];
  double x2[4], x3[4], g[4], q[4];
  double p, b11;
};

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

/* out[item][1280]: v1[256 rho + 64 mu + 16 nu + 4 lambda + sigma], then r1[64 mu + 16 nu + 4 lambda + sigma];
 * item = 3 * pair + evaluation; ev3 = 1 evaluates only (x, y) (pihat1 itself, item = pair) */
__global__ void __launch_bounds__(256) k_pihat1(size_t items, int ev3, const double *__restrict__ X,
                                                const double *__restrict__ Y, const signed char *__restrict__ tab,
                                                double *__restrict__ out)
{
  __shared__ PiConst C;
  const size_t item = blockIdx.x;
  const int t = threadIdx.x;
  const size_t pr = ev3 ? item / 3 : item;
  const int ev = ev3 ? (int)(item % 3) : 0;
  double x[4], y[4];
  {
    double px[4], py[4];
    for (int i = 0; i < 4; i++) { px[i] = X[4 * pr + i]; py[i] = Y[4 * pr + i]; }
    pert_args(ev, px, py, x, y);
  }
  const double pipi = M_PI * M_PI, twopipi = 2 * pipi;
  double xmy[4];
  for (int i = 0; i < 4; i++) xmy[i] = x[i] - y[i];
  double by[3];
  {
    const double yn = norm4(y);
    bessk2(yn, by);
    by[1] /= yn;
    by[2] /= (yn * yn);
  }
  if (t < 64) {
    const int c = t >> 4, d = (t >> 2) & 3, r = t & 3;
    C.fv[t] = pipi * (by[2] * y[c] * y[d] * y[r] + ((r == d) * y[c] - (c == r) * y[d] - (c == d) * y[r]) * by[1]);
  } else if (t < 128) {
    const int u = (t - 64) >> 4, r = (t >> 2) & 3, d = t & 3;
    const double p = twopipi * by[0];
    if (u == 0) { /* ell[r][d] = e[r] y[d] - delta s */
      const double e = twopipi * y[r] * by[2], s = twopipi * by[1];
      C.ell[4 * r + d] = r == d ? e * y[d] - s : e * y[d];
    } else if (u == 1) { /* h[r][d] = f[r] y[d] - delta p/2 */
      const double f = pipi * y[r] * by[1];
      C.h[4 * r + d] = r == d ? f * y[d] - 0.5 * p : f * y[d];
    } else if (u == 2) { /* fhat[d][r] = h[r][d] (+ p on the diagonal) */
      const double f = pipi * y[r] * by[1];
      const double hv = r == d ? f * y[d] - 0.5 * p : f * y[d];
      C.fhat[4 * d + r] = r == d ? hv + p : hv;
    } else {
      double bx[3], bz[3];
      const double xn = norm4(x), zn = norm4(xmy);
      bessk2(xn, bx);
      bx[1] /= xn;
      bx[2] /= (xn * xn);
      bessk2(zn, bz);
      bz[1] /= zn;
      bz[2] /= (zn * zn);
      const double b22 = bx[2] * bz[2];
      C.xab[4 * r + d] = x[r] * xmy[d] * b22;
      if (d == 0) {
        C.x2[r] = x[r] * (bx[2] * bz[1]);
        C.x3[r] = xmy[r] * (bx[1] * bz[2]);
        C.g[r] = pipi * y[r] * by[0];
        C.q[r] = 2 * (pipi * y[r] * by[1]);
      }
      if (r == 0 && d == 0) {
        C.p = p;
        C.b11 = bx[1] * bz[1];
      }
    }
  }
  __syncthreads();
  const int sg = t >> 6, lane = t & 63;
  const signed char *tm = tab + sg * 64 + lane;
  const signed char *t6 = tab + PT_MAIN + sg * 384 + lane;
  double t1[4] = { 0.0, 0.0, 0.0, 0.0 }, tr1 = 0.0;
  /* (alpha, beta) stay loops: the propagator factors stay in registers across them without spills */
#pragma unroll 1
  for (int a = 0; a < 4; a++) {
#pragma unroll 1
    for (int b = 0; b < 4; b++) {
      const int ab = 4 * a + b;
      const double xab = C.xab[ab];
      const signed char *gp = tm + ab * 4096;
      /* the reference's t1 -= fv * T as fused multiply-adds (its order; <= 2e-15 normwise apart) */
#pragma unroll
      for (int cd = 0; cd < 16; cd++) {
        const double T = xab * (double)gp[cd * 256];
        t1[0] = fma(-C.fv[4 * cd + 0], T, t1[0]);
        t1[1] = fma(-C.fv[4 * cd + 1], T, t1[1]);
        t1[2] = fma(-C.fv[4 * cd + 2], T, t1[2]);
        t1[3] = fma(-C.fv[4 * cd + 3], T, t1[3]);
        tr1 = fma(-C.ell[cd], T, tr1);
      }
      const signed char *sp = t6 + ab * 4 * 384;
      const double m1 = (double)sp[0] * xab;
      const double m2 = (double)sp[64] * C.x2[a] - (double)sp[128] * C.x3[a];
      const double m4 = (double)sp[192] * C.x2[a] - (double)sp[256] * C.x3[a];
      const double m6 = (double)sp[320] * C.b11;
#pragma unroll
      for (int r = 0; r < 4; r++)
        t1[r] += -C.g[r] * m1 - C.h[4 * r + b] * m2 - C.fhat[4 * b + r] * m4 + C.fv[16 * a + 4 * b + r] * m6;
      tr1 -= C.p * m1;
      tr1 -= C.q[b] * (m2 + m4);
      tr1 += C.ell[ab] * m6;
    }
  }
  const double m = (double)tab[PT_MAIN + PT_SIX + sg * 64 + lane] * C.b11;
  const double norm = 8.2336243479292e-07; /* 2 / (2 pi)^8, pi_pert.c:229 */
  double *o = out + item * 1280;
  const int mu = lane >> 4, nu = (lane >> 2) & 3, ld = lane & 3;
  const int ix = 64 * mu + 16 * nu + 4 * ld + sg;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    t1[r] += C.g[r] * m;
    o[256 * r + ix] = norm * t1[r];
  }
  tr1 += C.p * m;
  o[1024 + ix] = norm * tr1;
}

/* ipihatFermLoop_antisym (pi_pert.c:196-218) from the three pihat1 evaluations of each pair */
__global__ void __launch_bounds__(384) k_pihat_antisym(size_t n, const double *__restrict__ X, const double *__restrict__ V,
                                                       double *__restrict__ out)
{
  const size_t p = blockIdx.x;
  const int t = threadIdx.x;
  const int rs = t >> 6, mu = (t >> 4) & 3, nu = (t >> 2) & 3, ld = t & 3;
  const int RH[6] = { 0, 0, 0, 1, 1, 2 }, SG[6] = { 1, 2, 3, 2, 3, 3 };
  const int rho = RH[rs], sig = SG[rs];
  const double *v1 = V + 3 * p * 1280, *v2 = v1 + 1280, *v3 = v2 + 1280;
  const double *r2 = v2 + 1024, *r3 = v3 + 1024;
  const double xr = X[4 * p + rho], xs = X[4 * p + sig];
#define V5(A, a, b, c, d, e) A[256 * (a) + 64 * (b) + 16 * (c) + 4 * (d) + (e)]
#define R4(A, b, c, d, e) A[64 * (b) + 16 * (c) + 4 * (d) + (e)]
  out[p * 384 + t] =
      0.5 * (+V5(v1, rho, mu, nu, ld, sig) + V5(v2, rho, nu, ld, mu, sig) + xr * R4(r2, nu, ld, mu, sig) +
             V5(v3, rho, ld, nu, mu, sig) + xr * R4(r3, ld, nu, mu, sig) -
             (+V5(v1, sig, mu, nu, ld, rho) + V5(v2, sig, nu, ld, mu, rho) + xs * R4(r2, nu, ld, mu, rho) +
              V5(v3, sig, ld, nu, mu, rho) + xs * R4(r3, ld, nu, mu, rho)));
#undef V5
#undef R4
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

size_t kqed_pert_table_bytes(void) { return PT_MAIN + PT_SIX + PT_FOUR; }

void kqed_pert_build_table(const int *G8, signed char *tab)
{
  /* G8 index of (i0..i7) = sum_k i_k 4^(7-k); six- and four-gamma traces are G8 with leading 0, 0 */
  auto g8 = [G8](int a, int b, int c, int d, int e, int f, int g, int h) {
    return G8[h + 4 * (g + 4 * (f + 4 * (e + 4 * (d + 4 * (c + 4 * (b + 4 * a))))))];
  };
  for (int lane = 0; lane < 64; lane++) {
    const int mu = lane >> 4, nu = (lane >> 2) & 3, ld = lane & 3;
    for (int sg = 0; sg < 4; sg++) {
      for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) {
          for (int c = 0; c < 4; c++)
            for (int d = 0; d < 4; d++)
              tab[((((a * 4 + b) * 4 + c) * 4 + d) * 4 + sg) * 64 + lane] = (signed char)g8(a, mu, b, nu, c, sg, d, ld);
          signed char *s6 = tab + PT_MAIN + ((a * 4 + b) * 4 + sg) * 384 + lane;
          s6[0] = (signed char)g8(0, 0, a, mu, b, nu, sg, ld);
          s6[64] = (signed char)g8(0, 0, a, mu, nu, b, sg, ld);
          s6[128] = (signed char)g8(0, 0, mu, a, nu, b, sg, ld);
          s6[192] = (signed char)g8(0, 0, a, mu, nu, sg, b, ld);
          s6[256] = (signed char)g8(0, 0, mu, a, nu, sg, b, ld);
          s6[320] = (signed char)g8(0, 0, mu, nu, a, sg, b, ld);
        }
      tab[PT_MAIN + PT_SIX + sg * 64 + lane] = (signed char)g8(0, 0, 0, 0, mu, nu, sg, ld);
    }
  }
}

void kqed_pert_launch_pihat1(size_t n, const double *x, const double *y, const signed char *tab, double *out, hipStream_t s)
{
  hipLaunchKernelGGL(k_pihat1, dim3((unsigned)n), dim3(256), 0, s, n, 0, x, y, tab, out);
}

void kqed_pert_launch_ipihat(size_t n, const double *x, const double *y, const signed char *tab, double *scratch,
                             double *out, hipStream_t s)
{
  hipLaunchKernelGGL(k_pihat1, dim3((unsigned)(3 * n)), dim3(256), 0, s, 3 * n, 1, x, y, tab, scratch);
  hipLaunchKernelGGL(k_pihat_antisym, dim3((unsigned)n), dim3(384), 0, s, n, x, (const double *)scratch, out);
}

void kqed_pert_launch_scale(size_t n, double sc, const double *x, const double *y, double *xs, double *ys, hipStream_t s)
{
  hipLaunchKernelGGL(k_scale_pairs, dim3((unsigned)((4 * n + 255) / 256)), dim3(256), 0, s, n, sc, x, y, xs, ys);
}
