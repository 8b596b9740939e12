/*
 * kqed_combine.hip -- the reference's kernel combinations on the GPU.
 *
 * all_kernels.c, con_kernel.c, sub_kernel.c, SYMXY.c and SYMXY0.c build, per (x,y) pair, signed
 * sums of L0 at up to twelve derived argument pairs with permuted (mu,nu,lambda) indices (plus
 * the M-L2 Gaussian subtractions).  Here one pass of c pairs runs as
 *   k_comb_args   derived argument pairs, slot-major: pair i of slot j at j*c + i
 *   (pipeline)    L0 of all E*c derived pairs through the regular batched pipeline, AoS output
 *   k_combine     one 384-thread block per pair: the pair's E tensors staged in LDS, each thread
 *                 one (rhosig,mu,nu,lambda) of every output kernel, stores coalesced
 * The combination is a few FP64 adds per output and HBM-bound (E*3 KB read, NK*3 KB written per
 * pair); the L0 evaluations dominate.  Expressions keep the reference's operand order.
 */
#include "kqed_combine.h"

#include <hip/hip_runtime.h>

namespace {

/* derived argument pairs (x', y'); x-y = x_i - y_i, y-x = -(x-y), -y = -y_i, as the reference
 * forms them (con_kernel.c:66-67, all_kernels.c:296, SYMXY0.c:28-37) */
enum { CE_XY, CE_YX, CE_X_XMY, CE_XMY_X, CE_YMX_Y, CE_Y_YMX, CE_X_MY, CE_X0, CE_0X, CE_Y0, CE_0Y, CE_XMY0, CE_0XMY, CE_N };

__host__ __device__ constexpr unsigned comb_mask(int comb)
{
  constexpr unsigned ALL6 = (1u << CE_XY) | (1u << CE_YX) | (1u << CE_X0) | (1u << CE_0X) | (1u << CE_Y0) | (1u << CE_0Y);
  return comb == KQED_COMB_ALL || comb == KQED_COMB_ALL_M || comb == KQED_COMB_SYMXY || comb == KQED_COMB_SYMXY_V2 ? ALL6
       : comb == KQED_COMB_ALL_M_V2 ? ALL6 | (1u << CE_X_MY)
       : comb == KQED_COMB_CON || comb == KQED_COMB_CON_V2 ? ALL6 | (1u << CE_X_XMY) | (1u << CE_XMY0) | (1u << CE_0XMY)
       : comb == KQED_COMB_CON_M_L2 || comb == KQED_COMB_SUB_M_L2 ? ALL6 | (1u << CE_X_XMY) | (1u << CE_0XMY)
       : comb == KQED_COMB_SYMXY0 || comb == KQED_COMB_SYMXY0_V2
           ? ALL6 | (1u << CE_XMY_X) | (1u << CE_YMX_Y) | (1u << CE_X_XMY) | (1u << CE_Y_YMX) | (1u << CE_XMY0) | (1u << CE_0XMY)
           : 0u;
}

__host__ __device__ constexpr int comb_nk(int comb)
{
  return comb == KQED_COMB_CON_V2 || comb == KQED_COMB_CON_M_L2 || comb == KQED_COMB_SUB_M_L2 ? 12
       : comb == KQED_COMB_SYMXY_V2 || comb == KQED_COMB_SYMXY0_V2 ? 4
       : comb >= 0 && comb < KQED_COMB_N ? 8 : 0;
}

__host__ __device__ constexpr int popc(unsigned v) { return v ? (int)(v & 1u) + popc(v >> 1) : 0; }
/* compact slot of derived pair e within a combination's evaluations */
__host__ __device__ constexpr int slot(int comb, int e) { return popc(comb_mask(comb) & ((1u << e) - 1u)); }

__global__ void k_comb_args(size_t c, const double *__restrict__ x, const double *__restrict__ y, int ys, unsigned mask,
                            double *__restrict__ X, double *__restrict__ Y)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c) return;
  double xv[4], yv[4], xmy[4], ymx[4], my[4];
  for (int k = 0; k < 4; k++) {
    xv[k] = x[4 * i + k];
    yv[k] = y[(size_t)ys * i + k];
    xmy[k] = xv[k] - yv[k];
    ymx[k] = -xmy[k];
    my[k] = -yv[k];
  }
  int j = 0;
  for (int e = 0; e < CE_N; e++) {
    if (!(mask >> e & 1u)) continue;
    const double *a, *b;
    bool az = false, bz = false;
    switch (e) {
    case CE_XY: a = xv; b = yv; break;
    case CE_YX: a = yv; b = xv; break;
    case CE_X_XMY: a = xv; b = xmy; break;
    case CE_XMY_X: a = xmy; b = xv; break;
    case CE_YMX_Y: a = ymx; b = yv; break;
    case CE_Y_YMX: a = yv; b = ymx; break;
    case CE_X_MY: a = xv; b = my; break;
    case CE_X0: a = xv; b = xv; bz = true; break;
    case CE_0X: a = xv; b = xv; az = true; break;
    case CE_Y0: a = yv; b = yv; bz = true; break;
    case CE_0Y: a = yv; b = yv; az = true; break;
    case CE_XMY0: a = xmy; b = xmy; bz = true; break;
    default: a = xmy; b = xmy; az = true; break; /* CE_0XMY */
    }
    const size_t o = 4 * ((size_t)j * c + i);
    for (int k = 0; k < 4; k++) {
      X[o + k] = az ? 0.0 : a[k];
      Y[o + k] = bz ? 0.0 : b[k];
    }
    j++;
  }
}

/* The E staged L0 tensors of one pair and the thread's (rhosig, mu, nu, lambda). */
template <int COMB>
struct Pair {
  const double (*L)[384];
  int rs;
  __device__ double at(int e, int a, int b, int c) const { return L[slot(COMB, e)][64 * rs + 16 * a + 4 * b + c]; }
  /* sum_a v_a L_e with the contracted index in slot 1 or 2 of (mu,nu,lambda) */
  __device__ double dot1(const double *v, int e, int b, int c) const
  { return v[0] * at(e, 0, b, c) + v[1] * at(e, 1, b, c) + v[2] * at(e, 2, b, c) + v[3] * at(e, 3, b, c); }
  __device__ double dot2(const double *v, int e, int a, int c) const
  { return v[0] * at(e, a, 0, c) + v[1] * at(e, a, 1, c) + v[2] * at(e, a, 2, c) + v[3] * at(e, a, 3, c); }
};

/* compute_all_kernels (all_kernels.c:85-160): k[0..7] = L0.xy, L0.yx, ..., L3.yx at (m,n,l) */
template <int COMB>
__device__ void all_at(const Pair<COMB> &P, int m, int n, int l, double k[8])
{
  const double xml = P.at(CE_X0, l, n, m), yml = P.at(CE_Y0, l, n, m); /* i_to_mulam, :10-15 */
  const double f1 = (xml + yml) / 2.;
  const double f2 = P.at(CE_X0, m, n, l) + P.at(CE_0Y, m, n, l);
  const double f4 = P.at(CE_0X, m, n, l) + P.at(CE_Y0, m, n, l);
  const double f3 = P.at(CE_0X, m, n, l) - P.at(CE_0Y, m, n, l);
  const double a = P.at(CE_XY, m, n, l), b = P.at(CE_YX, m, n, l);
  k[0] = a;
  k[1] = b;
  k[2] = a + f1;
  k[3] = b + f1;
  k[4] = a - f2;
  k[5] = b - f4;
  k[6] = a + xml + f3;
  k[7] = b + yml - f3;
}

/* compute_all_kernels_SYMXY0_v2 (SYMXY0.c:21-106): k[0..3] at (m,n,l) */
template <int COMB>
__device__ void symxy0_at(const Pair<COMB> &P, int m, int n, int l, double k[4])
{
  const double k0 = (P.at(CE_XY, m, n, l) + P.at(CE_YX, n, m, l) - P.at(CE_XMY_X, n, l, m) - P.at(CE_YMX_Y, m, l, n) -
                     P.at(CE_X_XMY, l, n, m) - P.at(CE_Y_YMX, l, m, n)) / 6;
  const double apc = P.at(CE_X0, m, n, l) - P.at(CE_X0, l, n, m) + P.at(CE_Y0, n, m, l) - P.at(CE_Y0, l, m, n);
  const double k1 = k0 + (-apc + P.at(CE_X0, l, m, n) - P.at(CE_X0, m, l, n) + P.at(CE_Y0, l, n, m) - P.at(CE_Y0, n, l, m) +
                          P.at(CE_XMY0, n, m, l) - P.at(CE_XMY0, m, n, l) + P.at(CE_XMY0, n, l, m) - P.at(CE_XMY0, m, l, n)) / 12;
  const double k2 = k0 - (apc + P.at(CE_0X, n, m, l) - P.at(CE_0X, n, l, m) + P.at(CE_0Y, m, n, l) - P.at(CE_0Y, m, l, n) +
                          P.at(CE_XMY0, m, l, n) - P.at(CE_XMY0, n, l, m) + P.at(CE_0XMY, l, m, n) - P.at(CE_0XMY, l, n, m)) / 6.;
  const double k3 = k2 + (P.at(CE_0X, m, n, l) - P.at(CE_0X, l, n, m) + P.at(CE_0Y, n, m, l) - P.at(CE_0Y, l, m, n) +
                          P.at(CE_0XMY, m, l, n) - P.at(CE_0XMY, n, l, m)) / 6.;
  k[0] = k0;
  k[1] = k1;
  k[2] = k2;
  k[3] = k3;
}

struct CombArgs {
  double M[4];
};

/* DOT: instead of storing the outputs, dots[q * dstride + p] = sum_c K_q(x_p, y_p)_c Tc[p][c] (FP64,
 * fixed-order block reduction).  ys = 0 broadcasts one y to every pair. */
template <int COMB, bool DOT>
__global__ void __launch_bounds__(384) k_combine(size_t c, const double *__restrict__ L, const double *__restrict__ x,
                                                 const double *__restrict__ y, int ys, CombArgs A, double *__restrict__ out,
                                                 const float *__restrict__ Tc, size_t dstride)
{
  constexpr int E = popc(comb_mask(COMB));
  constexpr int NK = comb_nk(COMB);
  __shared__ double sh[E][384];
  __shared__ double g[12]; /* exp(-M_a v^2 / 2) for v = x, y, x-y */
  const size_t p = blockIdx.x;
  const int t = threadIdx.x;
  for (int j = 0; j < E; j++) sh[j][t] = L[((size_t)j * c + p) * 384 + t];
  double xv[4], yv[4], xmy[4];
  for (int k = 0; k < 4; k++) {
    xv[k] = x[4 * p + k];
    yv[k] = y[(size_t)ys * p + k];
    xmy[k] = xv[k] - yv[k];
  }
  if (t < 12) {
    const double *v = t < 4 ? xv : t < 8 ? yv : xmy;
    const double s = v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + v[3] * v[3];
    g[t] = exp(-A.M[t & 3] * s / 2.);
  }
  __syncthreads();
  const Pair<COMB> P = { sh, t / 64 };
  const int m = (t / 16) & 3, n = (t / 4) & 3, l = t & 3;
  const double *M = A.M;
  /* output q of this thread's (rhosig, mu, nu, lambda): stored, or multiplied by the site tensor */
  double acc[NK];
  double *o = DOT ? nullptr : out + p * (size_t)(NK * 384) + t;
  const double tv = DOT ? (double)Tc[p * 384 + t] : 0.0;
  auto put = [&](int q, double v) {
    if constexpr (DOT) acc[q] = v * tv;
    else o[384 * q] = v;
  };

  if constexpr (COMB == KQED_COMB_ALL) {
    double k[8];
    all_at(P, m, n, l, k);
    for (int q = 0; q < 8; q++) put(q, k[q]);
  } else if constexpr (COMB == KQED_COMB_SYMXY || COMB == KQED_COMB_SYMXY_V2) {
    /* SYMXY.c:12-85: (L^k.xy + L^k.yx_{nu mu lambda}) / 2, and for SYMXY .yx = .xy_{nu mu lambda} */
    double a[8], b[8];
    all_at(P, m, n, l, a);
    all_at(P, n, m, l, b);
    for (int q = 0; q < 4; q++) {
      const double xy = (a[2 * q] + b[2 * q + 1]) / 2;
      if constexpr (COMB == KQED_COMB_SYMXY_V2) {
        put(q, xy);
      } else {
        put((2 * q), xy);
        put((2 * q + 1), (b[2 * q] + a[2 * q + 1]) / 2);
      }
    }
  } else if constexpr (COMB == KQED_COMB_SYMXY0 || COMB == KQED_COMB_SYMXY0_V2) {
    double a[4];
    symxy0_at(P, m, n, l, a);
    if constexpr (COMB == KQED_COMB_SYMXY0_V2) {
      for (int q = 0; q < 4; q++) put(q, a[q]);
    } else {
      double b[4];
      symxy0_at(P, n, m, l, b);
      for (int q = 0; q < 4; q++) {
        put((2 * q), a[q]);
        put((2 * q + 1), b[q]);
      }
    }
  } else if constexpr (COMB == KQED_COMB_ALL_M || COMB == KQED_COMB_ALL_M_V2) {
    /* all_kernels.c:211-369: M-L2 subtractions with M[a] for kernel a */
    const double pxy1 = P.dot1(xv, CE_0Y, n, l), pxy2 = P.dot2(yv, CE_X0, m, l);
    for (int a = 0; a < 4; a++) {
      const double eX = g[a], eY = g[4 + a];
      const double sxy = eX * (P.at(CE_0Y, m, n, l) - M[a] * xv[m] * pxy1) + eY * (P.at(CE_X0, m, n, l) - M[a] * yv[n] * pxy2);
      if constexpr (COMB == KQED_COMB_ALL_M) {
        const double syx = eY * (P.at(CE_0X, m, n, l) - M[a] * yv[m] * P.dot1(yv, CE_0X, n, l)) +
                           eX * (P.at(CE_Y0, m, n, l) - M[a] * xv[n] * P.dot2(xv, CE_Y0, m, l));
        put((2 * a), P.at(CE_XY, m, n, l) - sxy);
        put((2 * a + 1), P.at(CE_YX, m, n, l) - syx);
      } else {
        const double syx2 = eY * (P.at(CE_0X, n, m, l) - M[a] * yv[n] * P.dot1(yv, CE_0X, m, l)) +
                            eX * (P.at(CE_Y0, n, m, l) - M[a] * xv[m] * P.dot2(xv, CE_Y0, n, l));
        const double sxmy = -eX * (P.at(CE_0Y, m, n, l) - M[a] * xv[m] * pxy1) + eY * (P.at(CE_X0, m, n, l) - M[a] * yv[n] * pxy2);
        put((2 * a), P.at(CE_XY, m, n, l) - sxy + P.at(CE_YX, n, m, l) - syx2);
        put((2 * a + 1), P.at(CE_X_MY, m, n, l) - sxmy);
      }
    }
  } else if constexpr (COMB == KQED_COMB_CON_V2 || COMB == KQED_COMB_CON) {
    /* con_kernel.c:61-244 */
    const double A0 = P.at(CE_XY, m, n, l), Yx = P.at(CE_YX, n, m, l), B0 = P.at(CE_X_XMY, l, n, m);
    const double C0 = A0 + Yx - B0;
    if constexpr (COMB == KQED_COMB_CON_V2) {
      const double A1 = A0 + (P.at(CE_X0, l, n, m) + P.at(CE_Y0, l, n, m)) / 2;
      const double B1 = B0 + (P.at(CE_X0, m, n, l) + P.at(CE_XMY0, m, n, l)) / 2;
      const double C1 = A1 - B1 + Yx + (P.at(CE_Y0, l, m, n) + P.at(CE_X0, l, m, n)) / 2;
      const double A2 = A0 - (P.at(CE_0Y, m, n, l) + P.at(CE_X0, m, n, l));
      const double B2 = B0 - (P.at(CE_0XMY, l, n, m) + P.at(CE_X0, l, n, m));
      const double C2 = A2 - B2 + Yx - (P.at(CE_0X, n, m, l) + P.at(CE_Y0, n, m, l));
      const double A3 = A0 + (P.at(CE_X0, l, n, m) + P.at(CE_0X, m, n, l) - P.at(CE_0Y, m, n, l));
      const double B3 = B0 + (P.at(CE_X0, m, n, l) + P.at(CE_0X, l, n, m) - P.at(CE_0XMY, l, n, m));
      const double C3 = A3 - B3 + Yx + (P.at(CE_Y0, l, m, n) + P.at(CE_0Y, n, m, l) - P.at(CE_0X, n, m, l));
      const double v[12] = { A0, A1, A2, A3, B0, B1, B2, B3, C0, C1, C2, C3 };
      for (int q = 0; q < 12; q++) put(q, v[q]);
    } else {
      const double c2 = C0 - (P.at(CE_0Y, m, n, l) + P.at(CE_X0, m, n, l) + P.at(CE_0X, n, m, l) + P.at(CE_Y0, n, m, l) -
                              P.at(CE_0XMY, l, n, m) - P.at(CE_X0, l, n, m));
      put(0, C0);
      put(1, B0);
      put(2, C0 + (P.at(CE_X0, l, n, m) + P.at(CE_Y0, l, n, m) + P.at(CE_X0, l, m, n) + P.at(CE_Y0, l, m, n) -
                         P.at(CE_X0, m, n, l) - P.at(CE_XMY0, m, n, l)) / 2);
      put(3, B0 + (P.at(CE_X0, m, n, l) + P.at(CE_XMY0, m, n, l)) / 2);
      put(4, c2);
      put(5, B0 - (P.at(CE_X0, l, n, m) + P.at(CE_0XMY, l, n, m)));
      put(6, c2 + (P.at(CE_0X, m, n, l) + P.at(CE_0Y, n, m, l) + P.at(CE_Y0, l, m, n) - P.at(CE_0X, l, n, m) +
                         P.at(CE_Y0, n, m, l)));
      put(7, B0 + (P.at(CE_X0, m, n, l) + P.at(CE_0X, l, n, m) - P.at(CE_0XMY, l, n, m)));
    }
  } else if constexpr (COMB == KQED_COMB_CON_M_L2 || COMB == KQED_COMB_SUB_M_L2) {
    /* con_kernel.c:246-399 (B read as lambda,nu,mu) and sub_kernel.c:26-167 (nu,lambda,mu) */
    constexpr bool SUB = COMB == KQED_COMB_SUB_M_L2;
    const int p1 = SUB ? n : l, p2 = SUB ? l : n; /* B = L(x,x-y)[p1][p2][mu] */
    const double pf1 = P.dot1(xv, CE_0Y, n, l), pf2 = P.dot2(yv, CE_X0, m, l);
    const double pf3 = P.dot1(yv, CE_0X, m, l), pf4 = P.dot2(xv, CE_Y0, n, l);
    const double pf5 = P.dot1(xv, CE_0XMY, p2, m), pf6 = P.dot2(xmy, CE_X0, p1, m);
    const double f5 = xv[p1], f6 = xmy[p2];
    for (int a = 0; a < 4; a++) {
      const double eX = g[a], eY = g[4 + a], eXMY = g[8 + a];
      const double lxy = P.at(CE_XY, m, n, l) - eX * (P.at(CE_0Y, m, n, l) - M[a] * xv[m] * pf1) -
                         eY * (P.at(CE_X0, m, n, l) - M[a] * yv[n] * pf2);
      const double lyx = P.at(CE_YX, n, m, l) - eY * (P.at(CE_0X, n, m, l) - M[a] * yv[n] * pf3) -
                         eX * (P.at(CE_Y0, n, m, l) - M[a] * xv[m] * pf4);
      const double lxm = P.at(CE_X_XMY, p1, p2, m) - eX * (P.at(CE_0XMY, p1, p2, m) - M[a] * f5 * pf5) -
                         eXMY * (P.at(CE_X0, p1, p2, m) - M[a] * f6 * pf6);
      put(a, lxy);
      put((4 + a), lxm);
      put((8 + a), lxy + lyx - lxm);
    }
  }
  if constexpr (DOT) {
    /* per output kernel: wavefront sums, then the six wavefront partials in order */
    __shared__ double part[6][NK];
    const int w = t >> 6, lane = t & 63;
#pragma unroll
    for (int q = 0; q < NK; q++) {
      double v = acc[q];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
      if (lane == 0) part[w][q] = v;
    }
    __syncthreads();
    if (t < NK) {
      double v = part[0][t];
      for (int j = 1; j < 6; j++) v += part[j][t];
      out[(size_t)t * dstride + p] = v;
    }
  }
}

template <int COMB>
void launch_combine(size_t c, const double *L, const double *x, const double *y, int ys, const double *M, double *out,
                    const float *Tc, size_t dstride, hipStream_t s)
{
  CombArgs A;
  for (int a = 0; a < 4; a++) A.M[a] = M ? M[a] : 0.0;
  if (Tc)
    hipLaunchKernelGGL((k_combine<COMB, true>), dim3((unsigned)c), dim3(384), 0, s, c, L, x, y, ys, A, out, Tc, dstride);
  else
    hipLaunchKernelGGL((k_combine<COMB, false>), dim3((unsigned)c), dim3(384), 0, s, c, L, x, y, ys, A, out, Tc, dstride);
}

} // namespace

int kqed_comb_evals(int comb) { return comb >= 0 && comb < KQED_COMB_N ? popc(comb_mask(comb)) : 0; }

extern "C" int kqed_comb_kernels(int comb) { return comb_nk(comb); }

void kqed_comb_launch_args(int comb, size_t c, const double *x, const double *y, int ys, double *X, double *Y, hipStream_t s)
{
  hipLaunchKernelGGL(k_comb_args, dim3((unsigned)((c + 255) / 256)), dim3(256), 0, s, c, x, y, ys, comb_mask(comb), X, Y);
}

void kqed_comb_launch(int comb, size_t c, const double *L, const double *x, const double *y, int ys, const double *M,
                      double *out, const float *Tc, size_t dstride, hipStream_t s)
{
  switch (comb) {
  case KQED_COMB_ALL: launch_combine<KQED_COMB_ALL>(c, L, x, y, ys, M, out, Tc, dstride, s); break;
  case KQED_COMB_ALL_M: launch_combine<KQED_COMB_ALL_M>(c, L, x, y, ys, M, out, Tc, dstride, s); break;
  case KQED_COMB_ALL_M_V2: launch_combine<KQED_COMB_ALL_M_V2>(c, L, x, y, ys, M, out, Tc, dstride, s); break;
  case KQED_COMB_CON: launch_combine<KQED_COMB_CON>(c, L, x, y, ys, M, out, Tc, dstride, s); break;
  case KQED_COMB_CON_V2: launch_combine<KQED_COMB_CON_V2>(c, L, x, y, ys, M, out, Tc, dstride, s); break;
  case KQED_COMB_CON_M_L2: launch_combine<KQED_COMB_CON_M_L2>(c, L, x, y, ys, M, out, Tc, dstride, s); break;
  case KQED_COMB_SUB_M_L2: launch_combine<KQED_COMB_SUB_M_L2>(c, L, x, y, ys, M, out, Tc, dstride, s); break;
  case KQED_COMB_SYMXY: launch_combine<KQED_COMB_SYMXY>(c, L, x, y, ys, M, out, Tc, dstride, s); break;
  case KQED_COMB_SYMXY_V2: launch_combine<KQED_COMB_SYMXY_V2>(c, L, x, y, ys, M, out, Tc, dstride, s); break;
  case KQED_COMB_SYMXY0: launch_combine<KQED_COMB_SYMXY0>(c, L, x, y, ys, M, out, Tc, dstride, s); break;
  case KQED_COMB_SYMXY0_V2: launch_combine<KQED_COMB_SYMXY0_V2>(c, L, x, y, ys, M, out, Tc, dstride, s); break;
  default: break;
  }
}
