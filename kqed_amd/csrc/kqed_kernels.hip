/*
 * kqed_kernels.hip -- HIP kernels (gfx950) for the position-space QED kernel L0..L3 / M-L2, FP64.
 *
 * Pipeline for a batch of (x,y) pairs (driven by kqed_api.hip):
 *   k_classify   regime + interpolation-cell key per pair, and its slot inside the cell
 *   k_classify / k_scan_tiles / k_scan_add / k_scatter   counting sort of the pairs by cell (table locality)
 *   k_interp     one thread per pair, in cell order: invariants, weight-function interpolation
 *                (Clenshaw sums over the precomputed per-node arrays), the chain-rule algebra of
 *                chnr_dS/dT/dV and the Taylor-table lerps; writes a per-pair record
 *   k_assemble   one block per group of pairs: S/T/V (or Taylor-table) tensors into LDS, then
 *                one thread per output component applies the Dirac-trace assembly map and the
 *                L1/L2/L3/M-L2 subtractions; coalesced stores
 * Built with -ffp-contract=off: every fused multiply-add below is written explicitly, where the
 * reference's AVX+FMA build fuses (simd.c:79-118, 424-435, 54-63; getff-new.c:69-90).
 */
#include <stdint.h>
#include <type_traits>

#include "../../include/kqed.h"
#include "kqed_device.h"

#define DEV __device__ __forceinline__
typedef kqed_dev_tables DT;

/* form-factor tables (reference KQED.h:52-54) */
enum { QG0, QG1, QG2, QG3, QL4, QL2, dxQG0, dxQG1, dxQG2, dxQG3, dxQL4, dxQL2, d2xQG2, d2xQG3 };
/* Taylor-x rows (reference KQED.h:57) */
enum { TYY = 0, G0dy, G0dx, Gl2, Gl2dy, Gl21, Gl21dy, Gl3, Gl3dy, G21, G21dy, G22A, G22B, G22Ady,
       G22Bdy, G3A, G3B, G3Ady, G3Bdy, G31A, G31B, G31Ady, G31Bdy };
/* Taylor-y rows (reference KQED.h:60) */
enum { alpha0dx_0p = 0, alpha0_1p, alpha3_0p, beta2_1p, alpha3_1p, alpha1_0p, beta4_1p, alpha1_1p,
       alpha1dx_1p, alpha3dx_0p, alpha3dxdx_0p, beta2dx_1p, alpha3dx_1p, alpha1dx_0p };

#define NYTAY 810
#define NXTAY 100

/* ------------------------------------------------------------------ small helpers */
DEV double lerp(double a, double t1, double t2) { return a * t1 + (1.0 - a) * t2; }
DEV double scal(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3]; }

/* y-power case of a table (reference getff-new.c:116-123): ypk 0: 1, 1: 1/y, 2: y, 3: 1/y^2 */
DEV void ffcase(int nm, int &ypk, int &mshm)
{
  switch (nm) {
  case QG2: case dxQG2: case d2xQG2: ypk = 1; mshm = 3; break;
  case QL4: case dxQL4: ypk = 2; mshm = 1; break;
  case QL2: case dxQL2: ypk = 3; mshm = 4; break;
  default: ypk = 0; mshm = 2; break;
  }
}
DEV int mm_of(int nm)
{
  int ypk, mshm;
  ffcase(nm, ypk, mshm);
  const int d = 2 - mshm;
  return d < 0 ? -d : d;
}

/* bisection of reference getff-new.c:199-210 */
DEV int bsrch(const double *__restrict__ arr, double target, int lo, int hi)
{
  while (hi - lo >= 2) {
    const int mid = (hi + lo) >> 1;
    if (arr[mid] > target) hi = mid; else lo = mid;
  }
  return lo;
}

struct Ip { double A, B, C1, C2, D, lA, y1, y2; int idx; };
struct Inv {
  double x, cb, y, xsq, ysq, xdoty, cborig, yorig, xmysq, xmy;
  bool flag2;
  Ip INVx, INVy;
};

/* precompute_INVx / precompute_INV (reference getff-new.c:30-56, simd.c:10-37) */
DEV void precomp(Ip &P, double y, double y1, double y2, int idx)
{
  P.idx = idx;
  const double dy = y1 - y2, ymy1 = y - y1, ymy2 = y - y2;
  const double ym2sq = ymy2 * ymy2, ym1sq = ymy1 * ymy1;
  P.A = -(ym2sq) * (2 * y - 3 * y1 + y2);
  P.B = (ym1sq) * (2 * y + y1 - 3 * y2);
  P.C1 = (ymy1) * (ym2sq) * dy;
  P.C2 = (ym1sq) * (ymy2) * dy;
  P.D = 1. / (dy * dy * dy);
  P.lA = (y2 - y) / (y2 - y1);
  P.y1 = y1;
  P.y2 = y2;
}

/* set_invariants (reference getff-new.c:316-370) */
DEV Inv set_inv(const DT &T, const double *xv, const double *yv)
{
  Inv I;
  const double EPSIN = 1E-7;
  I.xsq = scal(xv, xv); I.xsq = I.xsq < EPSIN ? EPSIN : I.xsq; I.x = sqrt(fabs(I.xsq));
  I.ysq = scal(yv, yv); I.ysq = I.ysq < EPSIN ? EPSIN : I.ysq; I.y = sqrt(fabs(I.ysq));
  I.xdoty = scal(xv, yv);
  I.cb = I.xdoty / (I.x * I.y);
  I.cborig = I.cb;
  I.yorig = I.y;
  I.xmysq = (I.xsq + I.ysq) * (1.00000000000001) - 2.0 * I.xdoty;
  I.xmy = sqrt(fabs(I.xmysq));
  I.flag2 = false;
  const double rx = (I.x > I.xmy ? I.xmy / I.x : I.x / I.xmy);
  const double rxy = (I.x > I.y ? I.y / I.x : I.x / I.y);
  if (rx < rxy && I.xmy < T.YY[T.nstpy - 1] && fabs(I.xmysq) > 1E-28) {
    I.cb = (I.x - I.y * I.cb) / I.xmy;
    I.y = I.xmy;
    I.flag2 = true;
  }
  const int ix1 = bsrch(T.XX, I.x, 0, T.nstpx);
  const int ix2 = ix1 + 1 >= T.nstpx ? ix1 : ix1 + 1;
  precomp(I.INVx, I.x, T.XX[ix1], T.XX[ix2], ix1);
  const int iy1 = bsrch(T.YY, I.y, 0, T.nstpy);
  const int iy2 = iy1 + 1 >= T.nstpy ? iy1 : iy1 + 1;
  precomp(I.INVy, I.y, T.YY[iy1], T.YY[iy2], iy1);
  return I;
}

DEV bool is_zero(const double *x) { return (x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + x[3] * x[3]) < 1E-28; }
DEV bool x_is_y(const double *x, const double *y)
{
  const double d[4] = { x[0] - y[0], x[1] - y[1], x[2] - y[2], x[3] - y[3] };
  return is_zero(d);
}

DEV size_t node(const DT &T, int nm, int ix, int iy) { return (((size_t)nm * T.nstpx + ix) * T.nstpy + iy) * T.ks; }

/* ------------------------------------------------------------------ table precompute */
/* One thread per (table, x node, y node): the per-node arrays the reference rebuilds for every
 * pair in getff6/getff7 (simd.c:343-613, FMA build) and getff2 (getff-new.c:104-156). */
__global__ void k_build_arrays(DT T, const float *__restrict__ Fm, const float *__restrict__ Fp,
                               double2 *__restrict__ P, double *__restrict__ PNC, double2 *__restrict__ S)
{
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nodes = (long)T.nstpx * T.nstpy;
  if (id >= 14 * nodes) return;
  const int nm = (int)(id / nodes);
  const int ix = (int)((id % nodes) / T.nstpy);
  const int iy = (int)(id % T.nstpy);
  const double y = T.YY[iy];
  const int nf = T.nfx[ix];
  int ypk, mshm;
  ffcase(nm, ypk, mshm);
  int mshp = 2 - mshm;
  const int mm = mshp < 0 ? -mshp : mshp;
  const float *fm = Fm + (((size_t)nm * T.nstpx + ix) * T.nstpy + iy) * T.nfmax;
  const float *fp = Fp + (((size_t)nm * T.nstpx + ix) * T.nstpy + iy) * T.nfmax;
  double2 *p = P + node(T, nm, ix, iy);
  double2 *s = S + node(T, nm, ix, iy);
  for (int k = 0; k < T.ks; k++) { p[k] = make_double2(0.0, 0.0); s[k] = make_double2(0.0, 0.0); }
  const double y1 = 1.0 / y;
  const double yp = ypk == 0 ? 1.0 : ypk == 1 ? y1 : ypk == 2 ? y : y1 * y1;
  /* packed path */
  {
    double facm = y1 * yp, facp = yp;
    facm = facm * y1;
    for (int j = 0; j < mm; j++) { facm = facm * y1; facp = facp * y; }
    double MSHM = mshm + mm, MSHP = mshp + mm;
    for (int j = 0; j < nf; j++) {
      const double vm = facm * (double)fm[j], vp = facp * (double)fp[j];
      double A = MSHM * vm;
      const double val = vm + vp;
      A = fma(MSHP, vp, -A);
      facm = facm * y1;
      facp = facp * y;
      MSHM = MSHM + 1.0;
      MSHP = MSHP + 1.0;
      p[mm + j] = make_double2(val, y1 * A);
    }
  }
  /* derivative-only variant (getff6 with calc = false, simd.c:452-475) */
  const int nci = nm == QG2 ? 0 : nm == dxQG2 ? 1 : nm == QG3 ? 2 : nm == dxQG3 ? 3 : -1;
  if (nci >= 0) {
    double *q = PNC + (((size_t)nci * T.nstpx + ix) * T.nstpy + iy) * T.ks;
    for (int k = 0; k < T.ks; k++) q[k] = 0.0;
    double facm = y1 * yp, facp = yp;
    facm = facm * y1;
    for (int j = 0; j < mm; j++) { facm = facm * y1; facp = facp * y; }
    double MSHM = mshm + mm, MSHP = mshp + mm;
    facm = facm * y1;
    facp = facp * y1;
    for (int j = 0; j < nf; j++) {
      const double vm = facm * (double)fm[j], vp = facp * (double)fp[j];
      const double A = MSHM * vm;
      facm = facm * y1;
      facp = facp * y;
      MSHM = MSHM + 1.0;
      q[mm + j] = fma(MSHP, vp, -A);
      MSHP = MSHP + 1.0;
    }
  }
  /* scalar path */
  {
    double facm = y1 * y1 * yp, facp = yp;
    for (int j = 0; j < mm; j++) { facm *= y1; facp *= y; }
    int im = mshm + mm, ip = mshp + mm;
    for (int j = 0; j < nf; j++) {
      const double dv = y1 * ip * facp * (double)fp[j] - y1 * im * facm * (double)fm[j];
      const double v = facm * (double)fm[j] + facp * (double)fp[j];
      s[mm + j] = make_double2(v, dv);
      facm *= y1; facp *= y; im++; ip++;
    }
  }
}

/* ------------------------------------------------------------------ Clenshaw sums */
/* Wave-uniform maximum of v over the active lanes (ballot/readlane; the loop runs once per distinct
 * value above the first lane's).  The result lives in a scalar register. */
DEV int wave_max_active(int v)
{
  int m = __builtin_amdgcn_readfirstlane(v);
  for (;;) {
    const unsigned long long gt = __ballot(v > m);
    if (!gt) return m;
    m = __builtin_amdgcn_readlane(v, __builtin_ctzll(gt));
  }
}

/* The ratio table read through the constant address space: with a wave-uniform index the loads
 * become scalar loads, so the per-step ratios cost no vector-memory bandwidth. */
typedef __attribute__((address_space(4))) const double cdouble;
DEV const cdouble *lut_ptr(const double *p) { return (const cdouble *)p; }

/* One packed-path Clenshaw step at index k (reference simd.c:79-118, 182-336, FMA build) for order d
 * on the chain pair (ya, yb); the ratios of order d > 0 come from the ratio table. */
template <int D>
DEV void clen_step(double &ya, double &yb, double c, double twox, double PA, double PB)
{
  const double t = yb;
  if (D == 0) {
    ya = ya - c;
    yb = fma(twox, yb, -ya);
  } else {
    /* PA is the step's ratio times 2 cos(beta), formed once per step for every chain of the order
     * (the reference multiplies yb by 2 cos(beta) first: the two forms agree to rounding) */
    ya = fma(PB, ya, -c);
    yb = fma(PA, yb, -ya);
  }
  ya = t;
}

/* Packed-path sums (reference simd.c:79-118, 182-336) over the arrays of two grid nodes of one x
 * node (y nodes iy and iy+1) for orders O0 .. O0+NO-1 of both the value (.x) and d/dy (.y) arrays
 * in one pass.  The reference evaluates sum_k c_k phi^d_k(cos beta) by Clenshaw's backward
 * recurrence, b_k = c_k + 2x PA^d_k b_{k+1} - PB^d_k b_{k+2} (PA, PB: the ratio table; order 0:
 * PA = PB = 1), result f_d b_d.  Here the same sums are formed directly: the polynomials come
 * from the forward recurrence that these coefficients define,
 *   phi_d = f_d, phi_{d-1} = 0, phi_{k+1} = 2x PA_k phi_k - PB_{k-1} phi_{k-1}   (f = 1, 2, 8, 48;
 * phi^1_k = U'_k, ...), one set per order shared by the four chains, and every chain is one
 * fused multiply-add per step into its own sum.  The two forms agree to rounding; the
 * accumulations are independent where Clenshaw's chains are serial, which the scheduler uses.
 * nk is wave-uniform (the largest length among the wave's nodes) and rounded up to whole blocks
 * of U steps: the arrays are zero past each node's own length (up to ks), so the extra steps add
 * exact zeros.  A block's U coefficient loads are issued together, one memory latency per block.
 * NT tables of one x node (equal lengths) run in one loop and share the polynomials. */
template <int O0, int NO, int NT, int U, typename P1>
DEV void clen_nodes(P1 const (&c)[NT], int ns, int nk, double twox, const double *__restrict__ LUTg,
                    double r[NT][2][NO][2])
{
  static_assert(U <= KQED_CLEN_BLOCK, "block longer than the table padding");
  const cdouble *LUT = lut_ptr(LUTg);
  double acc[NT][2][NO][2], ph[NO], pp[NO];
#pragma unroll
  for (int i = 0; i < NO; i++) {
    ph[i] = 0.0;
    pp[i] = 0.0;
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
      for (int a = 0; a < 2; a++) { acc[t][a][i][0] = acc[t][a][i][1] = 0.0; }
  }
  const int nsteps = nk - O0;
  const int kend = O0 + (nsteps > 0 ? (nsteps + U - 1) / U * U : 0);
  for (int kb = O0; kb < kend; kb += U) {
    double2 cc[U][NT][2];
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
      for (int t = 0; t < NT; t++) {
        /* double2 arrays (value, d/dy) read as pairs of doubles; node iy+1 is ns doubles on */
        cc[u][t][0] = make_double2(c[t][2 * (kb + u)], c[t][2 * (kb + u) + 1]);
        cc[u][t][1] = make_double2(c[t][ns + 2 * (kb + u)], c[t][ns + 2 * (kb + u) + 1]);
      }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int k = kb + u;
#pragma unroll
      for (int i = 0; i < NO; i++) {
        const int d = O0 + i;
        if (k < d) continue;
        if (k == d) {
          ph[i] = d == 0 ? 1.0 : d == 1 ? 2.0 : d == 2 ? 8.0 : 48.0;
          pp[i] = 0.0;
        } else {
          const double nw = d == 0 ? fma(twox, ph[i], -pp[i])
                                   : fma(twox * LUT[8 * (k - 1) + 2 * d - 1], ph[i], -(LUT[8 * (k - 2) + 2 * d] * pp[i]));
          pp[i] = ph[i];
          ph[i] = nw;
        }
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
          for (int a = 0; a < 2; a++) {
            acc[t][a][i][0] = fma(cc[u][t][a].x, ph[i], acc[t][a][i][0]);
            acc[t][a][i][1] = fma(cc[u][t][a].y, ph[i], acc[t][a][i][1]);
          }
      }
    }
  }
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int i = 0; i < NO; i++) { r[t][a][i][0] = acc[t][a][i][0]; r[t][a][i][1] = acc[t][a][i][1]; }
}

/* two packed chains of one order over plain double arrays (same wave-uniform length, blocks of U),
 * as direct sums over the polynomials of order d (see clen_nodes) */
template <int U, typename P1>
DEV void clen_two(P1 __restrict__ q1, P1 __restrict__ q2, int nk, int d, double twox,
                  const double *__restrict__ LUTg, double &o1, double &o2)
{
  static_assert(U <= KQED_CLEN_BLOCK, "block longer than the table padding");
  const cdouble *LUT = lut_ptr(LUTg);
  double s1 = 0.0, s2 = 0.0, pp = 0.0;
  double ph = d == 0 ? 1.0 : d == 1 ? 2.0 : d == 2 ? 8.0 : 48.0; /* phi_d */
  const int nsteps = nk - d;
  const int kend = d + (nsteps > 0 ? (nsteps + U - 1) / U * U : 0);
  for (int kb = d; kb < kend; kb += U) {
    double a[U], b[U];
#pragma unroll
    for (int u = 0; u < U; u++) { a[u] = q1[kb + u]; b[u] = q2[kb + u]; }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int k = kb + u;
      if (k > d) {
        const double nw = d == 0 ? fma(twox, ph, -pp)
                                 : fma(twox * LUT[8 * (k - 1) + 2 * d - 1], ph, -(LUT[8 * (k - 2) + 2 * d] * pp));
        pp = ph;
        ph = nw;
      }
      s1 = fma(a[u], ph, s1);
      s2 = fma(b[u], ph, s2);
    }
  }
  o1 = s1;
  o2 = s2;
}

/* scalar-path Clenshaw with divisions (reference cheby.c:13-91); comp selects .x or .y */
DEV double cheb_s(int order, int nk, double x, const double2 *__restrict__ co, int comp)
{
  double ya = 0.0, yb = 0.0, ytmp;
  const double twox = 2.0 * x;
  int j;
#define C(j) (comp ? co[j].y : co[j].x)
  switch (order) {
  case 0:
    for (j = nk - 1; j > 0; j--) { ytmp = yb; yb = twox * yb - ya + C(j); ya = ytmp; }
    return -ya + twox * yb + C(0);
  case 1:
    for (j = nk - 1; j > 0; j--) { ytmp = yb; yb = twox * (j + 1) * yb / j - (j + 3) * ya / (j + 1) + C(j); ya = ytmp; }
    return 2.0 * yb;
  case 2:
    for (j = nk - 1; j > 1; j--) {
      ytmp = yb; yb = twox * (j + 1) * yb / (j - 1); yb -= (j + 4) * ya / j; yb += C(j); ya = ytmp;
    }
    return 8 * yb;
  default:
    for (j = nk - 1; j > 2; j--) { ytmp = yb; yb = twox * (j + 1) * yb / (j - 2) - (j + 5) * ya / (j - 1) + C(j); ya = ytmp; }
    return 48 * yb;
  }
#undef C
}

/* ------------------------------------------------------------------ scalar interpolation path */
DEV double interpol3(double y, double y1, double y2, double dy, double f1, double f2, double g1, double g2)
{
  return ((-f1 * (y - y2) * (y - y2) * (2 * y - 3 * y1 + y2) +
           (y - y1) * (f2 * (y - y1) * (2 * y + y1 - 3 * y2) + (y - y2) * dy * (g1 * y + g2 * y - g2 * y1 - g1 * y2))) /
          (dy * dy * dy));
}

/* Scalar-path sums (reference cheby.c:13-91, which uses Clenshaw's recurrence with a division per
 * step) of order D over the value and d/dy arrays of one or two grid nodes in one pass:
 * r = {a.x, a.y, b.x, b.y}.  Formed directly over the polynomials of order D, phi^0_k = U_k,
 * phi^D_k from the forward recurrence with the ratio table (see clen_nodes): one set of
 * polynomials for the up-to-four chains, one fused multiply-add per chain and step (the forms
 * agree to rounding). */
template <int D, typename P1>
DEV void cheb_r(int nk, double x, P1 __restrict__ a, P1 __restrict__ b, const double *LUTg, double r[4])
{
  /* a, b: double2 arrays (value, d/dy) read as pairs of doubles */
  const cdouble *LUT = lut_ptr(LUTg);
  const int nc = b ? 4 : 2;
  double s[4] = { 0.0, 0.0, 0.0, 0.0 };
  const double twox = 2.0 * x;
  double ph = D == 0 ? 1.0 : D == 1 ? 2.0 : D == 2 ? 8.0 : 48.0, pp = 0.0; /* phi_D */
  /* steps in blocks of U with their loads issued together; steps past nk read the zero padding
   * of the arrays (ks slots, nk + U - 1 < ks) and add exact zeros */
  constexpr int U = 4;
  for (int j0 = D; j0 < nk; j0 += U) {
    double c[U][4];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int j = j0 + u;
      c[u][0] = a[2 * j]; c[u][1] = a[2 * j + 1];
      c[u][2] = b ? b[2 * j] : 0.0; c[u][3] = b ? b[2 * j + 1] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int j = j0 + u;
      if (j > D) {
        const double nw = D == 0 ? fma(twox, ph, -pp)
                                 : fma(twox * LUT[8 * (j - 1) + 2 * D - 1], ph, -(LUT[8 * (j - 2) + 2 * D] * pp));
        pp = ph;
        ph = nw;
      }
#pragma unroll
      for (int i = 0; i < 4; i++) {
        if (i >= nc) break;
        s[i] = fma(c[u][i], ph, s[i]);
      }
    }
  }
  for (int i = 0; i < 4; i++) r[i] = s[i];
}

/* getff2 + accessv (reference getff-new.c:104-194) over the precomputed scalar arrays, in two
 * parts: the Clenshaw sums of the value and d/dy arrays at y nodes iy, iy+1 (r = {a.x, a.y, b.x,
 * b.y}), which depend on (nm, ndcb) only, and their y interpolation.  Calls that differ only in
 * ndy, use_y_derivs or y share the sums. */
/* UNI: ix and iy are the same for every active lane of the wavefront (scalar registers), so the
 * coefficient arrays are read through the constant address space: scalar loads */
template <bool UNI = false>
DEV void sums_s(const DT &T, bool flag_hy, int ix, int iy, int nm, int ndcb, double cb, double r[4])
{
  const int mm = mm_of(nm);
  const int nk = T.nfx[ix] + mm;
  const int ord = mm + ndcb;
  if (UNI) {
    const cdouble *a = (const cdouble *)(T.S + node(T, nm, ix, iy));
    const cdouble *b = flag_hy ? nullptr : (const cdouble *)(T.S + node(T, nm, ix, iy + 1));
    switch (ord) {
    case 0: cheb_r<0>(nk, cb, a, b, T.LUT, r); break;
    case 1: cheb_r<1>(nk, cb, a, b, T.LUT, r); break;
    case 2: cheb_r<2>(nk, cb, a, b, T.LUT, r); break;
    default: cheb_r<3>(nk, cb, a, b, T.LUT, r); break;
    }
  } else {
    const double *a = (const double *)(T.S + node(T, nm, ix, iy));
    const double *b = flag_hy ? nullptr : (const double *)(T.S + node(T, nm, ix, iy + 1));
    switch (ord) {
    case 0: cheb_r<0>(nk, cb, a, b, T.LUT, r); break;
    case 1: cheb_r<1>(nk, cb, a, b, T.LUT, r); break;
    case 2: cheb_r<2>(nk, cb, a, b, T.LUT, r); break;
    default: cheb_r<3>(nk, cb, a, b, T.LUT, r); break;
    }
  }
}

/* sums_s for NT tables of equal mm at once (Taylor-x region: x node 0, y nodes iy and iy+1): one
 * set of polynomials for the 4 NT chains, the tables' coefficients loaded U steps at a time; each
 * chain's sum is sums_s's bit for bit (the extra steps past nk read zero padding) */
template <int D, int NT, int U, typename P1>
DEV void cheb_rn(int nk, double x, P1 const (&a)[NT], int ns, const double *LUTg, double r[NT][4])
{
  const cdouble *LUT = lut_ptr(LUTg);
  double s[NT][4];
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int i = 0; i < 4; i++) s[t][i] = 0.0;
  const double twox = 2.0 * x;
  double ph = D == 0 ? 1.0 : D == 1 ? 2.0 : D == 2 ? 8.0 : 48.0, pp = 0.0;
  for (int j0 = D; j0 < nk; j0 += U) {
    double c[U][NT][4];
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
      for (int t = 0; t < NT; t++) {
        const int j = j0 + u;
        c[u][t][0] = a[t][2 * j]; c[u][t][1] = a[t][2 * j + 1];
        c[u][t][2] = a[t][ns + 2 * j]; c[u][t][3] = a[t][ns + 2 * j + 1];
      }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int j = j0 + u;
      if (j > D) {
        const double nw = D == 0 ? fma(twox, ph, -pp)
                                 : fma(twox * LUT[8 * (j - 1) + 2 * D - 1], ph, -(LUT[8 * (j - 2) + 2 * D] * pp));
        pp = ph;
        ph = nw;
      }
#pragma unroll
      for (int t = 0; t < NT; t++)
#pragma unroll
        for (int i = 0; i < 4; i++) s[t][i] = fma(c[u][t][i], ph, s[t][i]);
    }
  }
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int i = 0; i < 4; i++) r[t][i] = s[t][i];
}

template <int NT, bool UNI>
DEV void tsums(const DT &T, const int (&nm)[NT], int ndcb, int iy, double cb, double r[NT][4])
{
  constexpr int U = NT == 1 ? 4 : 2;
  const int mm = mm_of(nm[0]);
  const int nk = T.nfx[0] + mm;
  const int ord = mm + ndcb;
  typedef typename std::conditional<UNI, const cdouble *, const double *>::type P1;
  P1 a[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) a[t] = (P1)(T.S + node(T, nm[t], 0, iy));
  const int ns = 2 * T.ks;
  switch (ord) {
  case 0: cheb_rn<0, NT, U>(nk, cb, a, ns, T.LUT, r); break;
  case 1: cheb_rn<1, NT, U>(nk, cb, a, ns, T.LUT, r); break;
  case 2: cheb_rn<2, NT, U>(nk, cb, a, ns, T.LUT, r); break;
  default: cheb_rn<3, NT, U>(nk, cb, a, ns, T.LUT, r); break;
  }
}

DEV double ysum_s(const DT &T, bool flag_hy, bool use_y_derivs, int iy, bool ndy, const double r[4], double y)
{
  const double r10 = ndy ? r[1] : r[0], r11 = r[1], r20 = ndy ? r[3] : r[2], r21 = r[3];
  if (!flag_hy) {
    const double y1 = T.YY[iy], y2 = T.YY[iy + 1];
    if (use_y_derivs) return interpol3(y, y1, y2, y1 - y2, r10, r20, r11, r21);
    return lerp((y2 - y) / (y2 - y1), r10, r20);
  }
  return r10;
}

template <bool UNI = false>
DEV double accessv_s(const DT &T, bool flag_hy, bool use_y_derivs, int ix, int iy, int nm, bool ndy, int ndcb,
                     double cb, double y)
{
  double r[4];
  sums_s<UNI>(T, flag_hy, ix, iy, nm, ndcb, cb, r);
  return ysum_s(T, flag_hy, use_y_derivs, iy, ndy, r, y);
}

DEV double interpol4(const Ip &P, double f1, double f2, double g1, double g2)
{
  return (f1 * P.A + f2 * P.B + g1 * P.C1 + g2 * P.C2) * P.D;
}

/* extractff on the scalar path (reference getff-new.c:236-257), used when a grid-edge flag is set */
DEV double extractff_s(const DT &T, int nm, bool ndy, int ndcb, const Inv &I)
{
  const bool flag_hx = (I.x >= T.XX[T.nstpx - 1]);
  const bool flag_hy = (I.y >= T.YY[T.nstpy - 1]);
  const bool use_x_derivs = (nm < dxQG0 || nm == dxQG2 || nm == dxQG3);
  const int ix = I.INVx.idx, iy = I.INVy.idx;
  const bool use_y_derivs = !ndy;
  const double f1iy = accessv_s(T, flag_hy, use_y_derivs, ix, iy, nm, ndy, ndcb, I.cb, I.y);
  if (!flag_hx) {
    const double f2iy = accessv_s(T, flag_hy, use_y_derivs, ix + 1, iy, nm, ndy, ndcb, I.cb, I.y);
    if (use_x_derivs) {
      const int offset = nm < dxQG0 ? dxQG0 : QL4;
      const double g1iy = accessv_s(T, flag_hy, use_y_derivs, ix, iy, nm + offset, ndy, ndcb, I.cb, I.y);
      const double g2iy = accessv_s(T, flag_hy, use_y_derivs, ix + 1, iy, nm + offset, ndy, ndcb, I.cb, I.y);
      return interpol4(I.INVx, f1iy, f2iy, g1iy, g2iy);
    }
    return lerp(I.INVx.lA, f1iy, f2iy);
  }
  return f1iy;
}

/* extractff2 on the scalar path (reference getff-new.c:278-292) */
DEV void extractff2_s(const DT &T, int nm, int ndcb, const Inv &I, double F[4])
{
  const int ndcb2 = ndcb + 1 < 5 ? ndcb + 1 : ndcb;
  const int dxmap[14] = { dxQG0, dxQG1, dxQG2, dxQG3, dxQL4, dxQL2, dxQG0, dxQG1, d2xQG2, d2xQG3, dxQL4, dxQL2, d2xQG2, d2xQG3 };
  F[0] = extractff_s(T, nm, false, ndcb2, I);
  F[1] = extractff_s(T, dxmap[nm], false, ndcb, I);
  F[2] = extractff_s(T, nm, true, ndcb, I);
  F[3] = extractff_s(T, nm, false, ndcb, I);
}

/* ------------------------------------------------------------------ packed interpolation path */
/* y step of one x node: INBLOCK (simd.c:54-63) and the d/dy lerp (simd.c:637-645) */
DEV double hy(const Ip &P, double v1, double v2, double d1, double d2)
{
  double t = v1 * P.A;
  t = fma(v2, P.B, t);
  t = fma(d1, P.C1, t);
  t = fma(d2, P.C2, t);
  return t * P.D;
}
DEV double ly(const Ip &P, double d1, double d2) { return fma(d1 - d2, P.lA, d2); }

/* interpol5 (getff-new.c:69-90): f = {H(nm,ix), H(nm,ix+1), H(dm,ix), H(dm,ix+1), L(nm,ix), L(nm,ix+1),
 * L(dm,ix), L(dm,ix+1), H'(nm,ix), H'(nm,ix+1), H'(dm,ix), H'(dm,ix+1)}, H' at order d+1. */
DEV void interpol5(double F[4], const Ip &P, const double f[12])
{
  double t;
  t = f[8] * P.A; t = fma(f[9], P.B, t); t = fma(f[10], P.C1, t); t = fma(f[11], P.C2, t); F[0] = t * P.D;
  const double ia = 1.0 / P.D, ib = P.lA / P.D, ic = -P.lA / P.D;
  t = f[3] * ia; t = fma(f[2], ib, t); t = fma(f[3], ic, t); t = t + 0.0; F[1] = t * P.D;
  t = f[4] * P.A; t = fma(f[5], P.B, t); t = fma(f[6], P.C1, t); t = fma(f[7], P.C2, t); F[2] = t * P.D;
  t = f[0] * P.A; t = fma(f[1], P.B, t); t = fma(f[2], P.C1, t); t = fma(f[3], P.C2, t); F[3] = t * P.D;
}

/* LDS of the packed-path kernel (dynamic, sized at launch): per wavefront KQED_LDS_ARRAYS node
 * arrays of ks slots, the staging area of its wave-uniform table reads */
extern __shared__ double2 kqed_interp_lds[];
DEV void wave_sync();

/* Per x node, the y-stepped values H (Hermite) and L (d/dy lerp) of NT tables of equal mm at NO
 * orders.  nfw: wave-uniform upper bound of nfx over the wave's pairs at this x-node position. */
template <int O0, int NO, int NT, int UNI>
DEV void table_xnodes(const DT &T, const int (&nm)[NT], int ixn, int nfw, const Inv &I, double twox, double H[NT][NO],
                      double Lv[NT][NO])
{
  /* steps per block: LDS-staged arrays (UNI) in blocks of 4, vector loads 8 (4 with more chains) */
  constexpr int U = UNI ? 4 : (NT * NO <= 2 ? 8 : 4);
  const int nk = nfw + mm_of(nm[0]);
  double r[NT][2][NO][2];
  if constexpr (UNI == 2) {
    /* two grid cells among the active lanes (A: the first lane's, B): the arrays of both are staged,
     * A's at arrays [0, 2NT), B's at [2NT, 4NT); each lane reads its own cell's (two addresses per
     * LDS read) */
    static_assert(4 * NT <= KQED_LDS_ARRAYS, "LDS staging area");
    const int iy = I.INVy.idx;
    const int ixa = __builtin_amdgcn_readfirstlane(ixn), iya = __builtin_amdgcn_readfirstlane(iy);
    const bool inb = ixn != ixa || iy != iya;
    const unsigned long long mb = __ballot(inb);
    const int fb = mb ? (int)__builtin_ctzll(mb) : 0;
    const int ixb = __builtin_amdgcn_readlane(ixn, fb), iyb = __builtin_amdgcn_readlane(iy, fb);
    const int nsteps = nk - O0;
    const int kend = O0 + (nsteps > 0 ? (nsteps + U - 1) / U * U : 0);
    double2 *wl = kqed_interp_lds + (threadIdx.x >> 6) * (KQED_LDS_ARRAYS * T.ks);
    const unsigned long long act = __ballot(1);
    const int na = __popcll(act);
    const int rank = __builtin_amdgcn_mbcnt_hi((unsigned)(act >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)act, 0));
    wave_sync(); /* the previous call's reads of the area come first */
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const double2 *ga = T.P + node(T, nm[t], ixa, iya), *gb = T.P + node(T, nm[t], ixb, iyb);
      for (int j = rank; j < kend; j += na) {
        wl[2 * t * T.ks + j] = ga[j];
        wl[(2 * t + 1) * T.ks + j] = ga[T.ks + j];
        wl[(2 * NT + 2 * t) * T.ks + j] = gb[j];
        wl[(2 * NT + 2 * t + 1) * T.ks + j] = gb[T.ks + j];
      }
    }
    wave_sync();
    const double *c[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) c[t] = (const double *)(wl + ((inb ? 2 * NT : 0) + 2 * t) * T.ks);
    clen_nodes<O0, NO, NT, U>(c, 2 * T.ks, nk, twox, T.LUT, r);
  } else if constexpr (UNI == 1) {
    /* one grid cell for every active lane: the wavefront copies the node arrays it needs (slots
     * [0, kend) of both y nodes of each table) into its LDS area with coalesced loads, then every
     * lane reads them there at the same address (broadcast), one LDS latency per block of steps */
    const int nsteps = nk - O0;
    const int kend = O0 + (nsteps > 0 ? (nsteps + U - 1) / U * U : 0);
    double2 *wl = kqed_interp_lds + (threadIdx.x >> 6) * (KQED_LDS_ARRAYS * T.ks);
    const unsigned long long act = __ballot(1);
    const int na = __popcll(act);
    const int rank = __builtin_amdgcn_mbcnt_hi((unsigned)(act >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)act, 0));
    wave_sync(); /* the previous call's reads of the area come first */
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const double2 *g = T.P + node(T, nm[t], ixn, I.INVy.idx);
      for (int j = rank; j < kend; j += na) {
        wl[2 * t * T.ks + j] = g[j];
        wl[(2 * t + 1) * T.ks + j] = g[T.ks + j];
      }
    }
    wave_sync();
    const double *c[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) c[t] = (const double *)(wl + 2 * t * T.ks);
    clen_nodes<O0, NO, NT, U>(c, 2 * T.ks, nk, twox, T.LUT, r);
  } else {
    const double *c[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) c[t] = (const double *)(T.P + node(T, nm[t], ixn, I.INVy.idx));
    clen_nodes<O0, NO, NT, U>(c, 2 * T.ks, nk, twox, T.LUT, r);
  }
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int i = 0; i < NO; i++) {
      H[t][i] = hy(I.INVy, r[t][0][i][0], r[t][1][i][0], r[t][0][i][1], r[t][1][i][1]);
      Lv[t][i] = ly(I.INVy, r[t][0][i][1], r[t][1][i][1]);
    }
}

template <int O0, int NO, int UNI>
DEV void table_xnode(const DT &T, int nm, int ixn, int nfw, const Inv &I, double twox, double H[NO], double Lv[NO])
{
  const int n1[1] = { nm };
  double H1[1][NO], L1[1][NO];
  table_xnodes<O0, NO, 1, UNI>(T, n1, ixn, nfw, I, twox, H1, L1);
#pragma unroll
  for (int i = 0; i < NO; i++) { H[i] = H1[0][i]; Lv[i] = L1[0][i]; }
}

/* derivative-only variant (accessv6 with ndy, simd.c:616-650) + x step: extractff(nm, ndy=true) */
template <bool UNI>
DEV double ndy_packed(const DT &T, int nci_nm, int nci_dm, int nm, const Inv &I, const int nfw[2], double twox, int d)
{
  double f[4];
  for (int xn = 0; xn < 2; xn++) {
    const int ixn = I.INVx.idx + xn;
    const int nk = nfw[xn] + mm_of(nm);
    for (int t = 0; t < 2; t++) {
      const int nci = t == 0 ? nci_nm : nci_dm;
      const double *q1 = T.PNC + (((size_t)nci * T.nstpx + ixn) * T.nstpy + I.INVy.idx) * T.ks;
      double a, b;
      if (UNI) {
        const cdouble *u1 = (const cdouble *)q1;
        clen_two<8>(u1, u1 + T.ks, nk, d, twox, T.LUT, a, b);
      } else {
        clen_two<8>(q1, q1 + T.ks, nk, d, twox, T.LUT, a, b);
      }
      f[2 * t + xn] = ly(I.INVy, a, b);
    }
  }
  /* QG2 and QG3 take the Hermite x step (use_x_derivs) */
  return interpol4(I.INVx, f[0], f[1], f[2], f[3]);
}

/* Two-table request (nm, dm) at orders d, d+1 assembled from per-x-node H/L values. */
DEV void request(const Ip &INVx, const double Hn[2][2], const double Ln[2], const double Hd[2][2], const double Ld[2],
                 double F[4])
{
  /* H?[xn][order offset 0/1], L?[xn] at order d */
  const double f[12] = { Hn[0][0], Hn[1][0], Hd[0][0], Hd[1][0], Ln[0], Ln[1], Ld[0], Ld[1],
                         Hn[0][1], Hn[1][1], Hd[0][1], Hd[1][1] };
  interpol5(F, INVx, f);
}

/* Simple groups: one request (nm, nm+6) at orders d, d+1 (d+1 < 4 always here). */
template <int D, int UNI>
DEV void simple_group(const DT &T, int nm, const Inv &I, const int nfw[2], double twox, double F[4])
{
  double Hn[2][2], Ln[2], Hd[2][2], Ld[2];
  const int nms[2] = { nm, nm + 6 }; /* a table and its x derivative: equal mm */
  for (int xn = 0; xn < 2; xn++) {
    double H[2][2], L[2][2];
    table_xnodes<D, 2, 2, UNI>(T, nms, I.INVx.idx + xn, nfw[xn], I, twox, H, L);
    Hn[xn][0] = H[0][0]; Hn[xn][1] = H[0][1]; Ln[xn] = L[0][0];
    Hd[xn][0] = H[1][0]; Hd[xn][1] = H[1][1]; Ld[xn] = L[1][0];
  }
  request(I.INVx, Hn, Ln, Hd, Ld, F);
}

/* ------------------------------------------------------------------ chain-rule algebra */
struct Tt {
  double ell1, ell2, ell3, v1, ell4, dv1dx, dv1dy, dv1dcb, dell2dx, dell2dy, dell2dcb;
  double dell1dx, dell1dy, dell1dcb, dell3dx, dell3dy, dell3dcb, dell4dy, dell4dx, dell4dcb;
};
struct Vt {
  double g2, dg1dx, dg1dy, dg1dcb, dg2dx, dg2dy, dg2dcb, dg3dx, dg3dy, dg3dcb;
  double ddg2dxdx, ddg2dxdy, ddg2dxdcb, ddg2dydcb, ddg2dcbdcb;
  double ddg3dxdx, ddg3dxdy, ddg3dxdcb, ddg3dydcb, ddg3dcbdcb;
  double ddg1dxdx, ddg1dxdy, ddg1dxdcb, ddg1dydcb, ddg1dcbdcb;
  double ddg1dydy, ddg2dydy;
};

/* chnr_dT standard branch algebra (reference chnr_dT.c:222-251) */
DEV void T_algebra(Tt &T, const Inv &I, double ysq)
{
  const double x = I.x, y = I.y, cb = I.cb, xsq = I.xsq;
  T.ell3 = T.ell4 / (2.0 * xsq * ysq) - y * cb * T.ell2 / x;
  T.ell1 = 4.0 / (3.0 * xsq * xsq) * (T.v1 - xsq * ysq * (cb * cb - 0.25) * T.ell2 - 1.5 * xsq * x * y * cb * T.ell3);
  T.dell3dy = (-T.ell4 / (ysq * y) + T.dell4dy / (2.0 * ysq) - cb * x * T.ell2 - y * x * cb * T.dell2dy) / (xsq);
  T.dell3dcb = (T.dell4dcb / (2.0 * x * ysq) - y * T.ell2 - y * cb * T.dell2dcb) / x;
  T.dell3dx = (-T.ell4 / (x * ysq) + T.dell4dx / (2.0 * ysq) + y * cb * T.ell2 - y * x * cb * T.dell2dx) / xsq;
  T.dell1dy = 4.0 / (3.0 * xsq * xsq) * (T.dv1dy - 2.0 * xsq * y * (cb * cb - 0.25) * T.ell2
                                         - xsq * ysq * (cb * cb - 0.25) * T.dell2dy
                                         - 1.5 * xsq * x * cb * T.ell3 - 1.5 * xsq * x * y * cb * T.dell3dy);
  T.dell1dcb = 4.0 / (3.0 * xsq * xsq) * (T.dv1dcb - 2.0 * xsq * ysq * cb * T.ell2
                                          - xsq * ysq * (cb * cb - 0.25) * T.dell2dcb
                                          - 1.5 * xsq * x * y * T.ell3 - 1.5 * xsq * x * y * cb * T.dell3dcb);
  T.dell1dx = -4.0 * T.ell1 / x
    + 4.0 / (3.0 * xsq * xsq) * (T.dv1dx - 2.0 * x * ysq * (cb * cb - 0.25) * T.ell2
                                 - xsq * ysq * (cb * cb - 0.25) * T.dell2dx
                                 - 4.5 * xsq * y * cb * T.ell3 - 1.5 * xsq * x * y * cb * T.dell3dx);
}

/* chnr_dV standard branch algebra (reference chnr_dV.c:413-442) */
DEV void V_algebra(Vt &D, const Inv &I)
{
  const double x = I.x, y = I.y, cb = I.cb, xsq = I.xsq;
  D.dg1dx = D.dg3dx + y / xsq * cb * D.g2 - y / x * cb * D.dg2dx;
  D.dg1dy = D.dg3dy - cb / x * D.g2 - (y / x) * cb * D.dg2dy;
  D.dg1dcb = D.dg3dcb - y / x * D.g2 - y / x * cb * D.dg2dcb;
  D.ddg1dxdx = D.ddg3dxdx - 2.0 * y / (xsq * x) * cb * D.g2 + y / xsq * cb * D.dg2dx + y / xsq * cb * D.dg2dx
    - y / x * cb * D.ddg2dxdx;
  D.ddg1dxdy = D.ddg3dxdy + cb / xsq * D.g2 + y * cb / xsq * D.dg2dy - cb / x * D.dg2dx - y / x * cb * D.ddg2dxdy;
  D.ddg1dxdcb = D.ddg3dxdcb + y / xsq * D.g2 + y / xsq * cb * D.dg2dcb - y / x * D.dg2dx - y / x * cb * D.ddg2dxdcb;
  D.ddg1dydcb = D.ddg3dydcb - D.g2 / x - cb / x * D.dg2dcb - y / x * D.dg2dy - y / x * cb * D.ddg2dydcb;
  D.ddg1dcbdcb = D.ddg3dcbdcb - 2.0 * y / x * D.dg2dcb - y / x * cb * D.ddg2dcbdcb;
}

/* Taylor-x branch of chnr_dS (reference chnr_dS.c:10-45) */
template <bool UNI>
DEV int S_taylor(const DT &T, double &dg0dy, double &dg0dx, double &dg0dcb, const Inv &I, int iy, int ity)
{
  const int iy_tay = ity;
  if (iy_tay == NYTAY - 1 || iy == T.nstpy - 1) return 1;
  const int i2 = iy_tay + 1;
  const double *TX = T.TX;
#define TXr(r, i) TX[(r) * NYTAY + (i)]
  const double ay = (TXr(TYY, i2) - I.y) / (TXr(TYY, i2) - TXr(TYY, iy_tay));
  const double ax = (T.XX[0] - I.x) / (T.XX[0]);
  double r0[2][4], r1[1][4];
  const int n0[2] = { dxQG0, QG0 }, n1[1] = { QG0 };
  tsums<2, UNI>(T, n0, 0, iy, I.cb, r0);
  tsums<1, UNI>(T, n1, 1, iy, I.cb, r1);
  double f1 = I.cb * lerp(ay, TXr(G0dx, iy_tay), TXr(G0dx, i2));
  double f2 = ysum_s(T, false, true, iy, false, r0[0], I.y);
  dg0dx = lerp(ax, f1, f2);
  f1 = lerp(ay, TXr(G0dy, iy_tay), TXr(G0dy, i2));
  f2 = ysum_s(T, false, false, iy, true, r0[1], I.y);
  dg0dy = lerp(ax, f1, f2);
  f2 = ysum_s(T, false, true, iy, false, r1[0], I.y);
  dg0dcb = lerp(ax, 0.0, f2);
  return 0;
}

/* Taylor-x branch of chnr_dT (reference chnr_dT.c:15-143) */
template <bool UNI>
DEV int T_taylor(const DT &T_, Tt &T, const Inv &I, int iy, int ity)
{
  const double ysq = I.flag2 ? I.xmysq : I.ysq;
  const int iy_tay = ity;
  if (iy_tay == NYTAY - 1 || iy == T_.nstpy - 1) return 1;
  const int i2 = iy_tay + 1;
  const double *TX = T_.TX;
  const double ay = (TXr(TYY, i2) - I.y) / (TXr(TYY, i2) - TXr(TYY, iy_tay));
  const double ax = (T_.XX[0] - I.x) / (T_.XX[0]);
  const double xb = T_.XX[0], xbsq = xb * xb, cb = I.cb, y = I.y;
  const double ell2a = lerp(ay, TXr(Gl2, iy_tay), TXr(Gl2, i2));
  /* node sums: a table and its x derivative (equal mm) in one loop; value and d/dy from the same sums */
  double r2[2][4], r1[1][4];
  const int nl2[2] = { QL2, dxQL2 }, ng1[2] = { QG1, dxQG1 }, nl4[2] = { QL4, dxQL4 };
  const int ol2[1] = { QL2 }, og1[1] = { QG1 }, ol4[1] = { QL4 };
  tsums<2, UNI>(T_, nl2, 0, iy, cb, r2);
  const double ell2b = ysum_s(T_, false, true, iy, false, r2[0], y);
  const double dell2bdy = ysum_s(T_, false, false, iy, true, r2[0], y);
  const double dell2bdx = ysum_s(T_, false, true, iy, false, r2[1], y);
  T.ell2 = lerp(ax, ell2a, ell2b);
  const double dell2adx = lerp(ay, TXr(Gl21, iy_tay), TXr(Gl21, i2)) * cb;
  T.dell2dx = lerp(ax, dell2adx, dell2bdx);
  const double dell2ady = lerp(ay, TXr(Gl2dy, iy_tay), TXr(Gl2dy, i2));
  T.dell2dy = lerp(ax, dell2ady, dell2bdy);
  tsums<1, UNI>(T_, ol2, 1, iy, cb, r1);
  const double dell2bdcb = ysum_s(T_, false, true, iy, false, r1[0], y);
  T.dell2dcb = lerp(ax, 0.0, dell2bdcb);
  tsums<2, UNI>(T_, ng1, 0, iy, cb, r2);
  const double v1b = ysum_s(T_, false, true, iy, false, r2[0], y);
  const double dv1bdy = ysum_s(T_, false, false, iy, true, r2[0], y);
  const double dv1bdx = ysum_s(T_, false, true, iy, false, r2[1], y);
  tsums<1, UNI>(T_, og1, 1, iy, cb, r1);
  const double dv1bdcb = ysum_s(T_, false, true, iy, false, r1[0], y);
  tsums<2, UNI>(T_, nl4, 0, iy, cb, r2);
  const double ell4b = ysum_s(T_, false, true, iy, false, r2[0], y);
  const double dell4bdy = ysum_s(T_, false, false, iy, true, r2[0], y);
  const double dell4bdx = ysum_s(T_, false, true, iy, false, r2[1], y);
  tsums<1, UNI>(T_, ol4, 1, iy, cb, r1);
  const double dell4bdcb = ysum_s(T_, false, true, iy, false, r1[0], y);
  const double ell3a = lerp(ay, TXr(Gl3, iy_tay), TXr(Gl3, i2));
  const double ell3b = ell4b / (2.0 * xbsq * ysq) - y * cb * ell2b / xb;
  T.ell3 = lerp(ax, ell3a, ell3b);
  const double ell1b = 4.0 / (3.0 * xbsq * xbsq) * (v1b - xbsq * ysq * (cb * cb - 0.25) * ell2b - 1.5 * xbsq * xb * y * cb * ell3b);
  T.ell1 = ell1b;
  const double dell3ady = lerp(ay, TXr(Gl3dy, iy_tay), TXr(Gl3dy, i2));
  const double dell3bdy = (-ell4b / (ysq * y) + dell4bdy / (2.0 * ysq) - cb * xb * ell2b - y * xb * cb * dell2bdy) / (xbsq);
  T.dell3dy = lerp(ax, dell3ady, dell3bdy);
  const double dell3bdcb = (dell4bdcb / (2.0 * xb * ysq) - y * ell2b - y * cb * dell2bdcb) / xb;
  T.dell3dcb = lerp(ax, 0.0, dell3bdcb);
  const double dell3bdx = (-ell4b / (xb * ysq) + dell4bdx / (2.0 * ysq) + y * cb * ell2b - y * xb * cb * dell2bdx) / xbsq;
  T.dell3dx = dell3bdx;
  T.dell1dy = 4.0 / (3.0 * xbsq * xbsq) * (dv1bdy - 2.0 * xbsq * y * (cb * cb - 0.25) * ell2b
                                           - xbsq * ysq * (cb * cb - 0.25) * dell2bdy
                                           - 1.5 * xbsq * xb * cb * ell3b - 1.5 * xbsq * xb * y * cb * dell3bdy);
  T.dell1dcb = 4.0 / (3.0 * xbsq * xbsq) * (dv1bdcb - 2.0 * xbsq * ysq * cb * ell2b - xbsq * ysq * (cb * cb - 0.25) * dell2bdcb
                                            - 1.5 * xbsq * xb * y * ell3b - 1.5 * xbsq * xb * y * cb * dell3bdcb);
  T.dell1dx = -4.0 * ell1b / xb + 4.0 / (3.0 * xbsq * xbsq)
    * (dv1bdx - 2.0 * xb * ysq * (cb * cb - 0.25) * ell2b - xbsq * ysq * (cb * cb - 0.25) * dell2bdx
       - 4.5 * xbsq * y * cb * ell3b - 1.5 * xbsq * xb * y * cb * dell3bdx);
  return 0;
}

/* Taylor-x branch of chnr_dV (reference chnr_dV.c:18-174) */
template <bool UNI>
DEV int V_taylor(const DT &T, Vt &D, double x, double y, double xsq, bool flag2, double cb, int iy, int ity)
{
  const int iy_tay = ity;
  if (iy_tay == NYTAY - 1 || iy == T.nstpy - 1) return 1;
  const int i2 = iy_tay + 1;
  const double *TX = T.TX;
  const double ay = (TXr(TYY, i2) - y) / (TXr(TYY, i2) - TXr(TYY, iy_tay));
  const double ax = (T.XX[0] - x) / (T.XX[0]);
  const double xb = T.XX[0], xbsq = xb * xb;
  /* calls that differ only in ndy (and, with flag2, in the interpolation point) share node sums */
  double r3[3][4], r2[2][4], r1[1][4];
  const int ng2[3] = { QG2, dxQG2, d2xQG2 }, ng2c[2] = { dxQG2, QG2 }, og2[1] = { QG2 };
  tsums<3, UNI>(T, ng2, 0, iy, cb, r3);
  const double g2b = ysum_s(T, false, true, iy, false, r3[0], y);
  const double dg2bdy = ysum_s(T, false, false, iy, true, r3[0], y);
  /* flag2: the same d/dy sums at the two y nodes (finite difference in y below) */
  const double dg2bdy_ya = flag2 ? ysum_s(T, false, false, iy, true, r3[0], T.YY[iy]) : 0.0;
  const double dg2bdy_yb = flag2 ? ysum_s(T, false, false, iy, true, r3[0], T.YY[iy + 1]) : 0.0;
  const double dg2bdx = ysum_s(T, false, true, iy, false, r3[1], y);
  const double ddg2bdxdy = ysum_s(T, false, false, iy, true, r3[1], y);
  const double ddg2bdxdx = ysum_s(T, false, true, iy, false, r3[2], y);
  tsums<2, UNI>(T, ng2c, 1, iy, cb, r2);
  const double ddg2bdxdcb = ysum_s(T, false, true, iy, false, r2[0], y);
  const double dg2bdcb = ysum_s(T, false, true, iy, false, r2[1], y);
  const double ddg2bdydcb = ysum_s(T, false, false, iy, true, r2[1], y);
  tsums<1, UNI>(T, og2, 2, iy, cb, r1);
  const double ddg2bdcbdcb = ysum_s(T, false, true, iy, false, r1[0], y);
  D.g2 = lerp(ax, 0.0, g2b);
  const double ddg2adxdcb = lerp(ay, TXr(G21, iy_tay), TXr(G21, i2));
  D.ddg2dxdcb = lerp(ax, ddg2adxdcb, ddg2bdxdcb);
  D.dg2dx = D.ddg2dxdcb * cb;
  D.dg2dcb = D.ddg2dxdcb * x;
  D.dg2dy = lerp(ax, 0.0, dg2bdy);
  const double ddg2adxdy = cb * lerp(ay, TXr(G21dy, iy_tay), TXr(G21dy, i2));
  D.ddg2dxdy = lerp(ax, ddg2adxdy, ddg2bdxdy);
  const double ddg2adxdx = 2.0 * (lerp(ay, TXr(G22A, iy_tay), TXr(G22A, i2))
                                  + lerp(ay, TXr(G22B, iy_tay), TXr(G22B, i2)) * (0.4 + 0.6 * (2.0 * cb * cb - 1)));
  D.ddg2dxdx = lerp(ax, ddg2adxdx, ddg2bdxdx);
  D.ddg2dcbdcb = ddg2bdcbdcb * xsq / xbsq;
  D.ddg2dydcb = lerp(ax, 0.0, ddg2bdydcb);
  const int ng3[3] = { QG3, dxQG3, d2xQG3 }, ng3c[2] = { dxQG3, QG3 }, og3[1] = { QG3 };
  tsums<3, UNI>(T, ng3, 0, iy, cb, r3);
  const double dg3bdy = ysum_s(T, false, false, iy, true, r3[0], y);
  const double dg3bdy_ya = flag2 ? ysum_s(T, false, false, iy, true, r3[0], T.YY[iy]) : 0.0;
  const double dg3bdy_yb = flag2 ? ysum_s(T, false, false, iy, true, r3[0], T.YY[iy + 1]) : 0.0;
  const double ddg3bdxdy = ysum_s(T, false, false, iy, true, r3[1], y);
  const double ddg3bdxdx = ysum_s(T, false, true, iy, false, r3[2], y);
  tsums<2, UNI>(T, ng3c, 1, iy, cb, r2);
  const double ddg3bdxdcb = ysum_s(T, false, true, iy, false, r2[0], y);
  const double ddg3bdydcb = ysum_s(T, false, false, iy, true, r2[1], y);
  tsums<1, UNI>(T, og3, 2, iy, cb, r1);
  const double ddg3bdcbdcb = ysum_s(T, false, true, iy, false, r1[0], y);
  const double dg1bdy = dg3bdy - cb / xb * g2b - (y / xb) * cb * dg2bdy;
  const double ddg1bdxdx = ddg3bdxdx - 2.0 * y / (xbsq * xb) * cb * g2b + y / xbsq * cb * dg2bdx + y / xbsq * cb * dg2bdx
    - y / xb * cb * ddg2bdxdx;
  const double ddg1bdxdy = ddg3bdxdy + cb / xbsq * g2b + y * cb / xbsq * dg2bdy - cb / xb * dg2bdx - y / xb * cb * ddg2bdxdy;
  const double ddg1bdxdcb = ddg3bdxdcb + y / xbsq * g2b + y / xbsq * cb * dg2bdcb - y / xb * dg2bdx - y / xb * cb * ddg2bdxdcb;
  const double ddg1bdydcb = ddg3bdydcb - g2b / xb - cb / xb * dg2bdcb - y / xb * dg2bdy - y / xb * cb * ddg2bdydcb;
  const double ddg1bdcbdcb = ddg3bdcbdcb - 2.0 * y / xb * dg2bdcb - y / xb * cb * ddg2bdcbdcb;
  const double dg1ady = lerp(ay, TXr(G3Ady, iy_tay), TXr(G3Ady, i2)) - lerp(ay, TXr(G3Bdy, iy_tay), TXr(G3Bdy, i2));
  const double ddg1adxdy = cb * (lerp(ay, TXr(G31Ady, iy_tay), TXr(G31Ady, i2))
                                 - (2 / 3.) * lerp(ay, TXr(G31Bdy, iy_tay), TXr(G31Bdy, i2))
                                 - lerp(ay, TXr(G22A, iy_tay), TXr(G22A, i2))
                                 - y * lerp(ay, TXr(G22Ady, iy_tay), TXr(G22Ady, i2)));
  const double ddg1adxdcb = lerp(ay, TXr(G31A, iy_tay), TXr(G31A, i2)) - (2 / 3.) * lerp(ay, TXr(G31B, iy_tay), TXr(G31B, i2))
    - y * lerp(ay, TXr(G22A, iy_tay), TXr(G22A, i2));
  const double ddg1adxdx = ddg1bdxdx;
  D.ddg1dxdcb = lerp(ax, ddg1adxdcb, ddg1bdxdcb);
  D.dg1dx = D.ddg1dxdcb * cb;
  D.dg1dcb = D.ddg1dxdcb * x;
  D.dg1dy = lerp(ax, dg1ady, dg1bdy);
  D.ddg1dxdy = lerp(ax, ddg1adxdy, ddg1bdxdy);
  D.ddg1dxdx = lerp(ax, ddg1adxdx, ddg1bdxdx);
  D.ddg1dcbdcb = ddg1bdcbdcb * xsq / xbsq;
  D.ddg1dydcb = lerp(ax, 0.0, ddg1bdydcb);
  if (flag2) {
    const double ddg2bdydy = (dg2bdy_yb - dg2bdy_ya) / (T.YY[1] - T.YY[0]);
    D.ddg2dydy = ax * 0.0 + (1.0 - ax) * ddg2bdydy;
    const double ddg1adydy = (TXr(G3Ady, i2) - TXr(G3Bdy, i2) - (TXr(G3Ady, iy_tay) - TXr(G3Bdy, iy_tay)))
      / (TXr(TYY, i2) - TXr(TYY, iy_tay));
    const double ddg3bdydy = (dg3bdy_yb - dg3bdy_ya) / (T.YY[1] - T.YY[0]);
    const double ddg1bdydy = ddg3bdydy - (y * ddg2bdydy + 2.0 * dg2bdy) * cb / x;
    D.ddg1dydy = ax * ddg1adydy + (1.0 - ax) * ddg1bdydy;
  }
  return 0;
}

/* x <-> x-y swap chain rules (reference chnr_dS.c:74-83, chnr_dT.c:146-180, chnr_dV.c:195-275) */
DEV void swap_S(double &dg0dx, double &dg0dcb, double &dg0dy, const Inv &I)
{
  const double dgs0dx = dg0dx, dgs0dcb = dg0dcb, dgs0dy = dg0dy;
  const double ca = I.cb;
  const double cd = (I.yorig - I.x * I.cborig) / I.xmy;
  dg0dx = dgs0dx + (1.0 - ca * ca) / I.xmy * dgs0dcb + ca * dgs0dy;
  dg0dcb = -(I.ysq * cd * dgs0dcb + I.x * I.yorig * I.xmy * dgs0dy) / I.xmysq;
  dg0dy = -(I.cborig + ca * cd) / I.xmy * dgs0dcb + cd * dgs0dy;
}

DEV void swap_T(Tt &T, const Inv &I)
{
  const double ell2s = T.ell2, dells2dx = T.dell2dx, dells2dcb = T.dell2dcb, dells2dy = T.dell2dy;
  const double cd = (I.yorig - I.x * I.cborig) / I.xmy, cb = I.cb;
  T.ell2 = ell2s;
  T.dell2dx = dells2dx + (1.0 - cb * cb) / I.xmy * dells2dcb + cb * dells2dy;
  T.dell2dcb = -(I.ysq * cd * dells2dcb + I.x * I.yorig * I.xmy * dells2dy) / I.xmysq;
  T.dell2dy = -(I.cborig + cb * cd) / I.xmy * dells2dcb + cd * dells2dy;
  const double ell3s = -T.ell3, dells3dx = -T.dell3dx, dells3dcb = -T.dell3dcb, dells3dy = -T.dell3dy;
  T.ell3 = ell3s - T.ell2;
  T.dell3dx = dells3dx + (1.0 - cb * cb) / I.xmy * dells3dcb + cb * dells3dy - T.dell2dx;
  T.dell3dcb = -(I.ysq * cd * dells3dcb + I.x * I.yorig * I.xmy * dells3dy) / I.xmysq - T.dell2dcb;
  T.dell3dy = -(I.cborig + cb * cd) / I.xmy * dells3dcb + cd * dells3dy - T.dell2dy;
  const double ell1s = T.ell1, dells1dx = T.dell1dx, dells1dcb = T.dell1dcb, dells1dy = T.dell1dy;
  T.ell1 = ell1s - T.ell2 - 2.0 * T.ell3;
  T.dell1dx = dells1dx + (1.0 - cb * cb) / I.xmy * dells1dcb + cb * dells1dy - T.dell2dx - 2.0 * T.dell3dx;
  T.dell1dcb = -(I.ysq * cd * dells1dcb + I.x * I.yorig * I.xmy * dells1dy) / I.xmysq - T.dell2dcb - 2.0 * T.dell3dcb;
  T.dell1dy = -(I.cborig + cb * cd) / I.xmy * dells1dcb + cd * dells1dy - T.dell2dy - 2.0 * T.dell3dy;
}

DEV void swap_V(Vt &D, double x, double xmy, double y, double xsq, double xmysq, double ysq, double ca, double cborig)
{
  Vt F = D;
  F.g2 = -D.g2; F.dg2dx = -D.dg2dx; F.dg2dy = -D.dg2dy; F.dg2dcb = -D.dg2dcb;
  F.ddg2dxdx = -D.ddg2dxdx; F.ddg2dxdy = -D.ddg2dxdy; F.ddg2dydy = -D.ddg2dydy;
  F.ddg2dxdcb = -D.ddg2dxdcb; F.ddg2dydcb = -D.ddg2dydcb; F.ddg2dcbdcb = -D.ddg2dcbdcb;
  const double cd = (y - x * cborig) / xmy;
  D.g2 = F.g2;
  D.dg2dx = F.dg2dx + (1.0 - ca * ca) / xmy * F.dg2dcb + ca * F.dg2dy;
  D.dg2dcb = -(ysq * cd * F.dg2dcb + x * y * xmy * F.dg2dy) / xmysq;
  D.dg2dy = -(cborig + ca * cd) / xmy * F.dg2dcb + cd * F.dg2dy;
  D.dg1dx = F.dg1dx + (1.0 - ca * ca) / xmy * F.dg1dcb + ca * F.dg1dy - D.dg2dx;
  D.dg1dcb = -(ysq * cd * F.dg1dcb + x * y * xmy * F.dg1dy) / xmysq - D.dg2dcb;
  D.dg1dy = -(cborig + ca * cd) / xmy * F.dg1dcb + cd * F.dg1dy - D.dg2dy;
  D.ddg2dxdx = (F.ddg2dxdx + (1.0 - ca * ca) / xmy * F.ddg2dxdcb + ca * F.ddg2dxdy)
    - 3.0 * ca * (1.0 - ca * ca) / xmysq * F.dg2dcb
    + (1.0 - ca * ca) / xmy * (F.ddg2dxdcb + (1.0 - ca * ca) / xmy * F.ddg2dcbdcb + ca * F.ddg2dydcb)
    + (1.0 - ca * ca) / xmy * F.dg2dy + ca * (F.ddg2dxdy + (1.0 - ca * ca) / xmy * F.ddg2dydcb + ca * F.ddg2dydy);
  D.ddg2dxdcb = ysq / (xmysq * xmy) * (cborig + 3.0 * ca * cd) * F.dg2dcb
    - ysq / xmysq * cd * (F.ddg2dxdcb + (1.0 - ca * ca) / xmy * F.ddg2dcbdcb + ca * F.ddg2dydcb)
    + y / xmysq * (x * ca - xmy) * F.dg2dy - x * y / xmy * (F.ddg2dxdy + (1.0 - ca * ca) / xmy * F.ddg2dydcb + ca * F.ddg2dydy);
  D.ddg2dxdy = (-(cborig + ca * cd) / xmy * F.ddg2dxdcb + cd * F.ddg2dxdy)
    + (2.0 * ca * cborig + cd * (3.0 * ca * ca - 1.0)) / xmysq * F.dg2dcb
    + (1.0 - ca * ca) / xmy * (-(cborig + ca * cd) / xmy * F.ddg2dcbdcb + cd * F.ddg2dydcb)
    - (cborig + ca * cd) / xmy * F.dg2dy + ca * (-(cborig + ca * cd) / xmy * F.ddg2dydcb + cd * F.ddg2dydy);
  D.ddg2dydcb = y / (xmysq * xmy) * ((3.0 * cd * cd - 1.0) * y - 2 * xmy * cd) * F.dg2dcb
    - ysq / (xmysq) * cd * (-(cborig + ca * cd) / xmy * F.ddg2dcbdcb + cd * F.ddg2dydcb)
    + x / xmysq * (y * cd - xmy) * F.dg2dy - x * y / xmy * (-(cborig + ca * cd) / xmy * F.ddg2dydcb + cd * F.ddg2dydy);
  D.ddg2dcbdcb = x * ysq / (xmysq * xmysq) * (xmy - 3.0 * y * cd) * F.dg2dcb
    - ysq / (xmysq) * cd * (-(ysq * cd * F.ddg2dcbdcb + x * y * xmy * F.ddg2dydcb) / xmysq)
    - xsq * ysq / (xmysq * xmy) * F.dg2dy
    - x * y / xmy * (-(ysq * cd * F.ddg2dydcb + x * y * xmy * F.ddg2dydy) / xmysq);
  D.ddg1dxdx = (F.ddg1dxdx + (1.0 - ca * ca) / xmy * F.ddg1dxdcb + ca * F.ddg1dxdy)
    - 3.0 * ca * (1.0 - ca * ca) / xmysq * F.dg1dcb
    + (1.0 - ca * ca) / xmy * (F.ddg1dxdcb + (1.0 - ca * ca) / xmy * F.ddg1dcbdcb + ca * F.ddg1dydcb)
    + (1.0 - ca * ca) / xmy * F.dg1dy + ca * (F.ddg1dxdy + (1.0 - ca * ca) / xmy * F.ddg1dydcb + ca * F.ddg1dydy)
    - D.ddg2dxdx;
  D.ddg1dxdcb = ysq / (xmysq * xmy) * (cborig + 3.0 * ca * cd) * F.dg1dcb
    - ysq / xmysq * cd * (F.ddg1dxdcb + (1.0 - ca * ca) / xmy * F.ddg1dcbdcb + ca * F.ddg1dydcb)
    + y / xmysq * (x * ca - xmy) * F.dg1dy
    - x * y / xmy * (F.ddg1dxdy + (1.0 - ca * ca) / xmy * F.ddg1dydcb + ca * F.ddg1dydy)
    - D.ddg2dxdcb;
  D.ddg1dxdy = (-(cborig + ca * cd) / xmy * F.ddg1dxdcb + cd * F.ddg1dxdy)
    + (2.0 * ca * cborig + cd * (3.0 * ca * ca - 1.0)) / xmysq * F.dg1dcb
    + (1.0 - ca * ca) / xmy * (-(cborig + ca * cd) / xmy * F.ddg1dcbdcb + cd * F.ddg1dydcb)
    - (cborig + ca * cd) / xmy * F.dg1dy + ca * (-(cborig + ca * cd) / xmy * F.ddg1dydcb + cd * F.ddg1dydy)
    - D.ddg2dxdy;
  D.ddg1dydcb = y / (xmysq * xmy) * ((3.0 * cd * cd - 1.0) * y - 2 * xmy * cd) * F.dg1dcb
    - ysq / (xmysq) * cd * (-(cborig + ca * cd) / xmy * F.ddg1dcbdcb + cd * F.ddg1dydcb)
    + x / xmysq * (y * cd - xmy) * F.dg1dy - x * y / xmy * (-(cborig + ca * cd) / xmy * F.ddg1dydcb + cd * F.ddg1dydy)
    - D.ddg2dydcb;
  D.ddg1dcbdcb = x * ysq / (xmysq * xmysq) * (xmy - 3.0 * y * cd) * F.dg1dcb
    - ysq / (xmysq) * cd * (-(ysq * cd * F.ddg1dcbdcb + x * y * xmy * F.ddg1dydcb) / xmysq)
    - xsq * ysq / (xmysq * xmy) * F.dg1dy
    - x * y / xmy * (-(ysq * cd * F.ddg1dydcb + x * y * xmy * F.ddg1dydy) / xmysq)
    - D.ddg2dcbdcb;
}

/* ------------------------------------------------------------------ standard-path evaluation */
/* Stage-1 record of a standard-regime pair: invariants and the interpolated weight functions
 * before the chain-rule algebra (the interpolation kernels write it field by field as each
 * table group finishes, so no group's results stay live in registers). */
struct Stage1 {
  double x, cb, y, xsq, ysq, xdoty, cborig, yorig, xmysq, xmy, flag2, need_alg;
  double dg0dx, dg0dy, dg0dcb;
  Tt T;
  Vt V;
};

/* the surrogate's slot list names these fields */
static_assert(offsetof(Stage1, dg0dcb) / 8 == 14 && offsetof(Stage1, dg0dx) / 8 == 12 && offsetof(Stage1, dg0dy) / 8 == 13,
              "stage-1 layout");
static_assert(offsetof(Stage1, T) / 8 + offsetof(Tt, dell2dx) / 8 == 23 && offsetof(Stage1, T) / 8 + offsetof(Tt, ell2) / 8 == 16 &&
              offsetof(Stage1, T) / 8 + offsetof(Tt, dell4dcb) / 8 == 34 && offsetof(Stage1, T) / 8 + offsetof(Tt, v1) / 8 == 18,
              "stage-1 layout");
static_assert(offsetof(Stage1, V) / 8 == 35 && offsetof(Stage1, V) / 8 + offsetof(Vt, dg3dy) / 8 == 43 &&
              offsetof(Stage1, V) / 8 + offsetof(Vt, ddg2dydy) / 8 == 61 &&
              offsetof(Stage1, V) / 8 + offsetof(Vt, ddg1dydy) / 8 == 60 &&
              offsetof(Stage1, V) / 8 + offsetof(Vt, ddg3dcbdcb) / 8 == 54 && offsetof(Stage1, flag2) / 8 == KQED_S1_FLAG2 &&
              offsetof(Stage1, need_alg) / 8 == KQED_S1_NEEDALG && sizeof(Stage1) <= 64 * 8,
              "stage-1 layout");

/* header of the stage-1 record (invariants) */
template <typename S1>
DEV void write_header(S1 *st, const Inv &I)
{
  st->x = I.x; st->cb = I.cb; st->y = I.y; st->xsq = I.xsq; st->ysq = I.ysq; st->xdoty = I.xdoty;
  st->cborig = I.cborig; st->yorig = I.yorig; st->xmysq = I.xmysq; st->xmy = I.xmy;
  st->flag2 = I.flag2 ? 1.0 : 0.0;
}

/* Surrogate bins: NB x NB over (log x, log y) of the post-swap invariants, x in [XX[0], XX[last]],
 * y in [KQED_SURR_YMIN, YY[last]] (clamped). */
DEV int surr_bin(const DT &T, double x, double y)
{
  const double lx0 = log(T.XX[0]), lx1 = log(T.XX[T.nstpx - 1]);
  const double ly0 = log(KQED_SURR_YMIN), ly1 = log(T.YY[T.nstpy - 1]);
  double u = (log(x) - lx0) / (lx1 - lx0), v = (log(y) - ly0) / (ly1 - ly0);
  int bu = (int)(u * KQED_SURR_NB), bv = (int)(v * KQED_SURR_NB);
  bu = bu < 0 ? 0 : bu >= KQED_SURR_NB ? KQED_SURR_NB - 1 : bu;
  bv = bv < 0 ? 0 : bv >= KQED_SURR_NB ? KQED_SURR_NB - 1 : bv;
  return bu * KQED_SURR_NB + bv;
}


/* The packed path (simd.c getff6/getff7 + accessv6/7 + extractff/extractff2, FMA build) for one
 * pair: every weight function of chnr_dS/dT/dV into st.  UNI: I's cell (INVx.idx, INVy.idx) and
 * nfw are the same in every active lane, so table addresses are scalar. */
template <int CLS, int UNI, typename S1>
DEV void packed_section(const DT &T, const Inv &I, S1 *__restrict__ st, const int nfw[2])
{
  auto &Q = st->T;
  auto &D = st->V;
  double F[4];
  const double twox = 2. * I.cb;
  /* chnr_dS: QG0 */
  simple_group<0, UNI>(T, QG0, I, nfw, twox, F);
  st->dg0dcb = F[0]; st->dg0dx = F[1]; st->dg0dy = F[2];
  /* chnr_dT: QL2, QL4, QG1 */
  simple_group<2, UNI>(T, QL2, I, nfw, twox, F);
  Q.dell2dx = F[1]; Q.dell2dy = F[2]; Q.ell2 = F[3]; Q.dell2dcb = F[0];
  simple_group<1, UNI>(T, QL4, I, nfw, twox, F);
  Q.dell4dx = F[1]; Q.dell4dy = F[2]; Q.ell4 = F[3]; Q.dell4dcb = F[0];
  simple_group<0, UNI>(T, QG1, I, nfw, twox, F);
  Q.dv1dx = F[1]; Q.dv1dy = F[2]; Q.v1 = F[3]; Q.dv1dcb = F[0];
  /* chnr_dV, QG2 family: tables QG2 (orders 1-3), dxQG2 (1-3), d2xQG2 (1-2) */
  {
    double H2[2][3], L2[2][3], H8[2][3], L8[2][3], H12[2][2], L12[2][2];
    const int n28[2] = { QG2, dxQG2 };
    for (int xn = 0; xn < 2; xn++) {
      double H[2][3], L[2][3];
      table_xnodes<1, 3, 2, UNI>(T, n28, I.INVx.idx + xn, nfw[xn], I, twox, H, L);
      for (int i = 0; i < 3; i++) { H2[xn][i] = H[0][i]; L2[xn][i] = L[0][i]; H8[xn][i] = H[1][i]; L8[xn][i] = L[1][i]; }
      table_xnode<1, 2, UNI>(T, d2xQG2, I.INVx.idx + xn, nfw[xn], I, twox, H12[xn], L12[xn]);
    }
    double Hn[2][2], Ln[2], Hd[2][2], Ld[2];
    /* extractff2(QG2, d0cb): d = 1 */
    for (int xn = 0; xn < 2; xn++) {
      Hn[xn][0] = H2[xn][0]; Hn[xn][1] = H2[xn][1]; Ln[xn] = L2[xn][0];
      Hd[xn][0] = H8[xn][0]; Hd[xn][1] = H8[xn][1]; Ld[xn] = L8[xn][0];
    }
    request(I.INVx, Hn, Ln, Hd, Ld, F);
    D.g2 = F[3]; D.dg2dy = F[2];
    /* extractff2(QG2, d1cb): d = 2 */
    for (int xn = 0; xn < 2; xn++) {
      Hn[xn][0] = H2[xn][1]; Hn[xn][1] = H2[xn][2]; Ln[xn] = L2[xn][1];
      Hd[xn][0] = H8[xn][1]; Hd[xn][1] = H8[xn][2]; Ld[xn] = L8[xn][1];
    }
    request(I.INVx, Hn, Ln, Hd, Ld, F);
    D.dg2dcb = F[3]; D.ddg2dydcb = F[2]; D.ddg2dcbdcb = F[0];
    /* extractff2(dxQG2, d0cb): d = 1 over (dxQG2, d2xQG2) */
    for (int xn = 0; xn < 2; xn++) {
      Hn[xn][0] = H8[xn][0]; Hn[xn][1] = H8[xn][1]; Ln[xn] = L8[xn][0];
      Hd[xn][0] = H12[xn][0]; Hd[xn][1] = H12[xn][1]; Ld[xn] = L12[xn][0];
    }
    request(I.INVx, Hn, Ln, Hd, Ld, F);
    D.ddg2dxdx = F[1]; D.dg2dx = F[3]; D.ddg2dxdy = F[2]; D.ddg2dxdcb = F[0];
  }
  /* chnr_dV, QG3 family: QG3 (orders 1-2), dxQG3 (0-2), d2xQG3 (0-1) */
  {
    double H3[2][2], L3[2][2], H9[2][3], L9[2][3], H13[2][2], L13[2][2];
    for (int xn = 0; xn < 2; xn++) {
      table_xnode<1, 2, UNI>(T, QG3, I.INVx.idx + xn, nfw[xn], I, twox, H3[xn], L3[xn]);
      table_xnode<0, 3, UNI>(T, dxQG3, I.INVx.idx + xn, nfw[xn], I, twox, H9[xn], L9[xn]);
      table_xnode<0, 2, UNI>(T, d2xQG3, I.INVx.idx + xn, nfw[xn], I, twox, H13[xn], L13[xn]);
    }
    double Hn[2][2], Ln[2], Hd[2][2], Ld[2];
    /* extractff2(dxQG3, d0cb): d = 0 over (dxQG3, d2xQG3) */
    for (int xn = 0; xn < 2; xn++) {
      Hn[xn][0] = H9[xn][0]; Hn[xn][1] = H9[xn][1]; Ln[xn] = L9[xn][0];
      Hd[xn][0] = H13[xn][0]; Hd[xn][1] = H13[xn][1]; Ld[xn] = L13[xn][0];
    }
    request(I.INVx, Hn, Ln, Hd, Ld, F);
    D.ddg3dxdx = F[1]; D.dg3dx = F[3]; D.ddg3dxdy = F[2]; D.ddg3dxdcb = F[0];
    /* extractff2(QG3, d1cb): d = 1 over (QG3, dxQG3) */
    for (int xn = 0; xn < 2; xn++) {
      Hn[xn][0] = H3[xn][0]; Hn[xn][1] = H3[xn][1]; Ln[xn] = L3[xn][0];
      Hd[xn][0] = H9[xn][1]; Hd[xn][1] = H9[xn][2]; Ld[xn] = L9[xn][1];
    }
    request(I.INVx, Hn, Ln, Hd, Ld, F);
    D.dg3dcb = F[3]; D.ddg3dydcb = F[2]; D.ddg3dcbdcb = F[0];
  }
  /* extractff(QG3, ndy, d0cb): derivative-only arrays (QG3 = 2, dxQG3 = 3) */
  const double dg3dy = ndy_packed<UNI == 1>(T, 2, 3, QG3, I, nfw, twox, 0);
  D.dg3dy = dg3dy;
  if (I.flag2) {
    /* reference chnr_dV.c:444-466: both finite-difference ends reuse the stencil weights of y,
     * so they coincide unless the upper end sits on the top y node */
    const int iy1 = I.INVy.idx;
    const double fq1 = ndy_packed<UNI == 1>(T, 0, 1, QG2, I, nfw, twox, 1);
    double fq2 = fq1, gq2 = dg3dy;
    if (CLS == 0 && iy1 + 1 == T.nstpy - 1) {
      Inv I2 = I;
      I2.y = T.YY[iy1 + 1];
      fq2 = extractff_s(T, QG2, true, 0, I2);
      gq2 = extractff_s(T, QG3, true, 0, I2);
    }
    const double ddg2dydy = (fq2 - fq1) / (T.YY[1] - T.YY[0]);
    D.ddg2dydy = ddg2dydy;
    const double ddg3dydy = (gq2 - dg3dy) / (T.YY[1] - T.YY[0]);
    D.ddg1dydy = ddg3dydy - (I.y * ddg2dydy + 2.0 * D.dg2dy) * I.cb / I.x;
  }
}

/* Interpolation of one pair into st.  CLS: 0 any pair, 1 packed-path pairs only, 2 Taylor-x pairs
 * only (the sort sends each class to its own kernel, which then carries only its own code path).
 * Returns 1 when the pair is outside the grid (the reference then leaves an all-zero kernel). */
/* private (lane-interleaved) memory: a field's stores and loads across a wavefront are coalesced */
typedef __attribute__((address_space(5))) double PD;
template <typename MP>
DEV void algebra_finish(MP m, Stage1 &s, const double *hx, const double *hy);
DEV void algebra_wave(double (*t)[17], double *m, Stage1 &s, const double *hx, const double *hy);

template <int CLS>
DEV int interp_eval(const DT &T, const double *xv, const double *yv, Stage1 *__restrict__ st, bool fuse, bool &alg_done,
                    double (*tw)[17])
{
  const Inv I = set_inv(T, xv, yv);
  if (CLS != 1 && (CLS == 2 || I.x < T.XX[0])) {
    /* Taylor-x region: every chnr_* takes its Taylor branch (scalar path, x node 0); the Taylor
     * branches already apply the algebra (chnr_dT.c:16-143, chnr_dV.c:19-174) */
    double dg0dx, dg0dy, dg0dcb;
    Tt Q;
    for (int i = 0; i < (int)(sizeof(Tt) / 8); i++) ((double *)&Q)[i] = 0.0;
    Vt D;
    for (int i = 0; i < (int)(sizeof(Vt) / 8); i++) ((double *)&D)[i] = 0.0;
    const int iy = bsrch(T.YY, I.y, 0, T.nstpy);
    /* the sort keys the Taylor region by y node, so a wavefront's pairs nearly always share it:
     * then every coefficient array is read by scalar loads */
    const int iyu = __builtin_amdgcn_readfirstlane(iy);
    const int ity = bsrch(T.TX + TYY * NYTAY, I.y, 0, NYTAY); /* the Taylor table's y row, once for S, T, V */
    if (CLS == 2 && __ballot(iy != iyu) == 0) {
      if (S_taylor<true>(T, dg0dy, dg0dx, dg0dcb, I, iyu, ity)) return 1;
      if (T_taylor<true>(T, Q, I, iyu, ity)) return 1;
      if (V_taylor<true>(T, D, I.x, I.y, I.xsq, I.flag2, I.cb, iyu, ity)) return 1;
    } else {
      if (S_taylor<false>(T, dg0dy, dg0dx, dg0dcb, I, iy, ity)) return 1;
      if (T_taylor<false>(T, Q, I, iy, ity)) return 1;
      if (V_taylor<false>(T, D, I.x, I.y, I.xsq, I.flag2, I.cb, iy, ity)) return 1;
    }
    if (fuse) {
      /* the chain-rule algebra straight from registers: the stage-1 record is never written */
      Stage1 s1;
      write_header(&s1, I);
      s1.need_alg = 0.0;
      s1.dg0dx = dg0dx; s1.dg0dy = dg0dy; s1.dg0dcb = dg0dcb;
      s1.T = Q;
      s1.V = D;
      /* the packed-path and Taylor-region kernels store through their LDS tile */
      if constexpr (CLS == 0) algebra_finish((double *)st, s1, xv, yv);
      else algebra_wave(tw, (double *)st, s1, xv, yv);
      alg_done = true;
      return 0;
    }
    write_header(st, I);
    st->dg0dx = dg0dx; st->dg0dy = dg0dy; st->dg0dcb = dg0dcb;
    st->T = Q;
    st->V = D;
    st->need_alg = 0.0;
    return 0;
  }
  if (CLS == 2) return 1;
  if constexpr (CLS == 1) {
    /* the stage-1 values go to private memory, whose lane-interleaved layout makes every store and
     * load of a field one coalesced access (the record's fields lie 1152 bytes apart from lane to
     * lane); the algebra then reads them back from there */
    typedef __attribute__((address_space(5))) Stage1 PS1;
    Stage1 loc;
    unsigned pa = (unsigned)(size_t)(PS1 *)&loc;
    asm volatile("" : "+v"(pa)); /* an opaque address: the fields stay in memory, not registers */
    PS1 *ps = (PS1 *)(size_t)pa;
    write_header(ps, I);
    ps->need_alg = 1.0;
    const int nfw[2] = { wave_max_active(T.nfx[I.INVx.idx]), wave_max_active(T.nfx[I.INVx.idx + 1]) };
    const int cell = I.INVx.idx * T.nstpy + I.INVy.idx;
    const int cu = __builtin_amdgcn_readfirstlane(cell);
    const unsigned long long other = __ballot(cell != cu);
    if (other == 0) {
      Inv Iu = I;
      Iu.INVx.idx = __builtin_amdgcn_readfirstlane(I.INVx.idx);
      Iu.INVy.idx = __builtin_amdgcn_readfirstlane(I.INVy.idx);
      packed_section<CLS, 1>(T, Iu, ps, nfw);
    } else {
      const int c2 = __builtin_amdgcn_readlane(cell, (int)__builtin_ctzll(other));
      if (__ballot(cell != cu && cell != c2) == 0) packed_section<CLS, 2>(T, I, ps, nfw);
      else packed_section<CLS, 0>(T, I, ps, nfw);
    }
    const PD *pd = (const PD *)ps;
    if (fuse) {
      Stage1 s1;
      for (int i = 0; i < (int)(sizeof(Stage1) / 8); i++) ((double *)&s1)[i] = pd[i];
      algebra_wave(tw, (double *)st, s1, xv, yv);
      alg_done = true;
    } else {
      for (int i = 0; i < (int)(sizeof(Stage1) / 8); i++) ((double *)st)[i] = pd[i];
    }
    return 0;
  }
  if (CLS != 1 && I.x > T.XX[T.nstpx - 1]) return 1;
  write_header(st, I);
  st->need_alg = 1.0;
  Tt &Q = st->T;
  Vt &D = st->V;
  const bool flag_hx = (I.x >= T.XX[T.nstpx - 1]);
  const bool flag_hy = (I.y >= T.YY[T.nstpy - 1]);
  double F[4];
  if (CLS == 1 || (!flag_hx && !flag_hy)) {
    /* sum lengths are made wave-uniform (zero padding keeps each node's sums exact), so the
     * loops and ratio-table reads are scalar */
    const int nfw[2] = { wave_max_active(T.nfx[I.INVx.idx]), wave_max_active(T.nfx[I.INVx.idx + 1]) };
    /* the sort keys the packed path by cell, so most wavefronts hold one cell only: then the
     * cell's indices are wave-uniform and every table read is a scalar load (the other
     * wavefronts read through vector loads) */
    const int cell = I.INVx.idx * T.nstpy + I.INVy.idx;
    const int cu = __builtin_amdgcn_readfirstlane(cell);
    bool uni = false;
    if constexpr (CLS == 1) {
      const unsigned long long other = __ballot(cell != cu);
      if (other == 0) {
        Inv Iu = I;
        Iu.INVx.idx = __builtin_amdgcn_readfirstlane(I.INVx.idx);
        Iu.INVy.idx = __builtin_amdgcn_readfirstlane(I.INVy.idx);
        packed_section<CLS, 1>(T, Iu, st, nfw);
        uni = true;
      } else {
        /* two cells (the sorted order's cell boundary inside the wavefront): both staged in LDS */
        const int c2 = __builtin_amdgcn_readlane(cell, (int)__builtin_ctzll(other));
        if (__ballot(cell != cu && cell != c2) == 0) {
          packed_section<CLS, 2>(T, I, st, nfw);
          uni = true;
        }
      }
    }
    if (!uni) packed_section<CLS, 0>(T, I, st, nfw);
  } else if (CLS == 0) {
    /* grid-edge fallback: every extractff2 on the scalar path */
    extractff2_s(T, QG0, 0, I, F); st->dg0dcb = F[0]; st->dg0dx = F[1]; st->dg0dy = F[2];
    extractff2_s(T, QL2, 0, I, F); Q.dell2dx = F[1]; Q.dell2dy = F[2]; Q.ell2 = F[3]; Q.dell2dcb = F[0];
    extractff2_s(T, QL4, 0, I, F); Q.dell4dx = F[1]; Q.dell4dy = F[2]; Q.ell4 = F[3]; Q.dell4dcb = F[0];
    extractff2_s(T, QG1, 0, I, F); Q.dv1dx = F[1]; Q.dv1dy = F[2]; Q.v1 = F[3]; Q.dv1dcb = F[0];
    D.dg3dy = extractff_s(T, QG3, true, 0, I);
    extractff2_s(T, QG2, 0, I, F); D.g2 = F[3]; D.dg2dy = F[2];
    extractff2_s(T, dxQG2, 0, I, F); D.ddg2dxdx = F[1]; D.dg2dx = F[3]; D.ddg2dxdy = F[2]; D.ddg2dxdcb = F[0];
    extractff2_s(T, dxQG3, 0, I, F); D.ddg3dxdx = F[1]; D.dg3dx = F[3]; D.ddg3dxdy = F[2]; D.ddg3dxdcb = F[0];
    extractff2_s(T, QG2, 1, I, F); D.dg2dcb = F[3]; D.ddg2dydcb = F[2]; D.ddg2dcbdcb = F[0];
    extractff2_s(T, QG3, 1, I, F); D.dg3dcb = F[3]; D.ddg3dydcb = F[2]; D.ddg3dcbdcb = F[0];
    if (I.flag2) {
      const int iy1 = bsrch(T.YY, I.y, 0, T.nstpy);
      if (iy1 == T.nstpy - 1) return 1;
      Inv I1 = I, I2 = I;
      I1.y = T.YY[iy1];
      I2.y = T.YY[iy1 + 1];
      double fq1 = extractff_s(T, QG2, true, 0, I1), fq2 = extractff_s(T, QG2, true, 0, I2);
      const double ddg2dydy = (fq2 - fq1) / (T.YY[1] - T.YY[0]);
      D.ddg2dydy = ddg2dydy;
      fq1 = extractff_s(T, QG3, true, 0, I1);
      fq2 = extractff_s(T, QG3, true, 0, I2);
      const double ddg3dydy = (fq2 - fq1) / (T.YY[1] - T.YY[0]);
      D.ddg1dydy = ddg3dydy - (I.y * ddg2dydy + 2.0 * D.dg2dy) * I.cb / I.x;
    }
  }
  return 0;
}

/* Chain-rule algebra of chnr_dT/chnr_dV (standard branch), the x <-> x-y swap, and the 36
 * primitives the assembly kernel reads (S: 3, T: 12, invariants: 6, V: 15), written in place. */
DEV void algebra_values(Stage1 &s, const double *hx, const double *hy, double o[64])
{
  Inv I;
  I.x = s.x; I.cb = s.cb; I.y = s.y; I.xsq = s.xsq; I.ysq = s.ysq; I.xdoty = s.xdoty;
  I.cborig = s.cborig; I.yorig = s.yorig; I.xmysq = s.xmysq; I.xmy = s.xmy; I.flag2 = s.flag2 != 0.0;
  if (s.need_alg != 0.0) {
    T_algebra(s.T, I, I.flag2 ? I.xmysq : I.ysq);
    V_algebra(s.V, I);
  }
  if (I.flag2) {
    swap_S(s.dg0dx, s.dg0dcb, s.dg0dy, I);
    swap_T(s.T, I);
    swap_V(s.V, I.x, I.xmy, I.yorig, I.xsq, I.xmysq, I.ysq, I.cb, I.cborig);
  }
  const Tt &Q = s.T;
  const Vt &D = s.V;
  o[0] = s.dg0dx; o[1] = s.dg0dy; o[2] = s.dg0dcb;
  o[3] = Q.ell1; o[4] = Q.ell2; o[5] = Q.ell3;
  o[6] = Q.dell1dx; o[7] = Q.dell1dy; o[8] = Q.dell1dcb;
  o[9] = Q.dell2dx; o[10] = Q.dell2dy; o[11] = Q.dell2dcb;
  o[12] = Q.dell3dx; o[13] = Q.dell3dy; o[14] = Q.dell3dcb;
  o[15] = I.x; o[16] = I.yorig; o[17] = I.cborig; o[18] = I.xsq; o[19] = I.ysq; o[20] = I.xdoty;
  o[21] = D.dg1dx; o[22] = D.dg2dx; o[23] = D.dg1dcb; o[24] = D.dg2dcb; o[25] = D.dg1dy;
  o[26] = D.ddg1dxdx; o[27] = D.ddg2dxdx; o[28] = D.ddg1dxdy; o[29] = D.ddg2dxdy;
  o[30] = D.ddg1dxdcb; o[31] = D.ddg2dxdcb; o[32] = D.ddg1dydcb; o[33] = D.ddg2dydcb;
  o[34] = D.ddg1dcbdcb; o[35] = D.ddg2dcbdcb;
  /* Per-pair vectors and sums the assembly's helper vectors are built from (std_helper), written
   * after the primitives: xh = x/|x|, yh = y/|y|, c2 = (xh - cb yh)/|y|, c4 = (yh - cb xh)/|x|,
   * e12 = xh/|y| + yh/|x| - cb (xh/|x| + yh/|y|), P21 + P22, P23 + P24, 1/|x|, 1/|y|, 1, 0. */
  const double rx = 1.0 / I.x, ry = 1.0 / I.yorig, cb = I.cborig;
  double *h = o + KQED_HELP_BASE;
  for (int i = 0; i < 4; i++) {
    const double xh = hx[i] * rx, yh = hy[i] * ry;
    h[i] = xh;
    h[4 + i] = yh;
    h[8 + i] = (xh - cb * yh) * ry;
    h[12 + i] = (yh - cb * xh) * rx;
    h[16 + i] = xh * ry + yh * rx - cb * (xh * rx + yh * ry);
  }
  h[20] = o[21] + o[22];
  h[21] = o[23] + o[24];
  h[22] = rx;
  h[23] = ry;
  h[24] = 1.0;
  h[25] = 0.0;
  o[62] = o[63] = 0.0;
}

template <typename MP>
DEV void algebra_finish(MP m, Stage1 &s, const double *hx, const double *hy)
{
  double o[64];
  algebra_values(s, hx, hy, o);
  for (int i = 0; i < 36; i++) m[i] = o[i];
  for (int i = 0; i < 26; i++) m[KQED_HELP_BASE + i] = o[KQED_HELP_BASE + i];
}

/* Wavefront-wide store of 16 doubles per lane: lane l's v[0..15] become doubles 0..15 of the
 * record at base + l * REC_LEN (base: lane 0's record, or the same offset in it), for the lanes in
 * mask m.  The values pass through the wavefront's LDS tile t[64][17], so each store instruction
 * writes 128 contiguous bytes of each of four records, where the lanes' own stores would each touch
 * a different line (the records lie 1152 bytes apart).  The whole wavefront calls it together (lane l
 * stores field l & 15 of records l >> 4, 4 + (l >> 4), ...). */
DEV void wave_store16(double (*t)[17], double *base, const double *v, unsigned long long m)
{
  const int lane = __lane_id();
  for (int k = 0; k < 16; k++) t[lane][k] = v[k];
  wave_sync();
  for (int j = 0; j < 16; j++) {
    const int rr = 4 * j + (lane >> 4);
    if ((m >> rr) & 1) base[(size_t)rr * KQED_REC_LEN + (lane & 15)] = t[rr][lane & 15];
  }
  wave_sync();
}

/* algebra_finish with the outputs of a whole wavefront stored together through its LDS tile t
 * (wave_store16); a wavefront with lanes missing (the last one of a segment, pairs past the Taylor
 * table) stores each lane's own record instead */
DEV void algebra_wave(double (*t)[17], double *m, Stage1 &s, const double *hx, const double *hy)
{
  double o[64];
  algebra_values(s, hx, hy, o);
  if (__ballot(1) == ~0ull) {
    double *base = m - __lane_id() * KQED_REC_LEN;
    for (int c = 0; c < 4; c++) wave_store16(t, base + 16 * c, o + 16 * c, ~0ull);
  } else {
    for (int i = 0; i < 62; i++) m[i] = o[i];
  }
}

DEV void algebra_main(double *__restrict__ m, const double *hx, const double *hy)
{
  Stage1 s = *(const Stage1 *)m;
  algebra_finish(m, s, hx, hy);
}

/* The algebra as its own pass (records whose stage-1 values come from elsewhere: the surrogate) */
__global__ void __launch_bounds__(256)
k_algebra(size_t n, double *__restrict__ rec, const unsigned *__restrict__ kinds)
{
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  /* only the records whose stage-1 values came from elsewhere: the interpolation kernels ran the
   * algebra of every other record themselves */
  if (p >= n || (kinds[p] & 7u) != (unsigned)KIND_STD || !(kinds[p] & KQED_SURR_BIT)) return;
  double *r = rec + p * KQED_REC_LEN;
  algebra_main(r + KQED_REC_MAIN, r + KQED_REC_X, r + KQED_REC_Y);
}

/* Caller-supplied weight functions into the stage-1 records of the packed-path pairs the
 * interpolation kernel flagged (KQED_SURR_BIT): w[p][34] for input pair p, the 34 in the order of
 * KQED_SURR_SLOTS except that the last two are the second y derivatives of g2 and g3 (swap regime
 * only; ddg1dydy follows from them as in the interpolation, chnr_dV.c:466). */
__global__ void __launch_bounds__(256)
k_inject_records(size_t n, double *__restrict__ rec, const unsigned *__restrict__ kinds, const double *__restrict__ w)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !(kinds[i] & KQED_SURR_BIT)) return;
  double *r = rec + i * KQED_REC_LEN;
  double *m = r + KQED_REC_MAIN;
  const double *src = w + (size_t)r[KQED_REC_P] * KQED_NSURR;
  const int slots[KQED_NSURR] = KQED_SURR_SLOTS;
  for (int k = 0; k < KQED_NSURR - 2; k++) m[slots[k]] = src[k];
  Stage1 *st = (Stage1 *)m;
  if (st->flag2 != 0.0) {
    const double ddg2dydy = src[KQED_NSURR - 2], ddg3dydy = src[KQED_NSURR - 1];
    st->V.ddg2dydy = ddg2dydy;
    st->V.ddg1dydy = ddg3dydy - (st->y * ddg2dydy + 2.0 * st->V.dg2dy) * st->cb / st->x;
  } else {
    st->V.ddg2dydy = 0.0;
    st->V.ddg1dydy = 0.0;
  }
}

/* ------------------------------------------------------------------ Taylor-table primitives */
/* Tabd_xeq0 lerps (reference Tabd.c:23-69): out = v[4], y, ysq, fx, fy, ell2, dell2adx, dell2ady,
 * ell3a, dg2dx, ddg2dxdy, dg1dx, dg1dy, phi1, phi2.  Returns 1 when |v| is past the table. */
DEV int xeq0_prims(const DT &T, const double *yv, double *o)
{
  const double ysq = scal(yv, yv);
  const double y = sqrt(fabs(ysq));
  const int i1 = bsrch(T.TX + TYY * NYTAY, y, 0, NYTAY);
  if (i1 == NYTAY - 1) return 1;
  const int i2 = i1 + 1;
  const double *TX = T.TX;
  const double ay = (TXr(TYY, i2) - y) / (TXr(TYY, i2) - TXr(TYY, i1));
  for (int i = 0; i < 4; i++) o[i] = yv[i];
  o[4] = y; o[5] = ysq;
  o[6] = 0.5 * lerp(ay, TXr(G0dx, i1), TXr(G0dx, i2));
  o[7] = 0.5 * lerp(ay, TXr(G0dy, i1), TXr(G0dy, i2));
  o[8] = lerp(ay, TXr(Gl2, i1), TXr(Gl2, i2));
  o[9] = lerp(ay, TXr(Gl21, i1), TXr(Gl21, i2));
  o[10] = lerp(ay, TXr(Gl2dy, i1), TXr(Gl2dy, i2));
  o[11] = lerp(ay, TXr(Gl3, i1), TXr(Gl3, i2));
  o[12] = lerp(ay, TXr(G21, i1), TXr(G21, i2));
  o[13] = lerp(ay, TXr(G21dy, i1), TXr(G21dy, i2));
  o[14] = (lerp(ay, TXr(G31A, i1), TXr(G31A, i2)) - (2. / 3) * (lerp(ay, TXr(G31B, i1), TXr(G31B, i2)))
           - y * (lerp(ay, TXr(G22A, i1), TXr(G22A, i2))));
  o[15] = (lerp(ay, TXr(G3Ady, i1), TXr(G3Ady, i2)) - lerp(ay, TXr(G3Bdy, i1), TXr(G3Bdy, i2)));
  o[16] = (lerp(ay, TXr(G22A, i1), TXr(G22A, i2)) - lerp(ay, TXr(G22B, i1), TXr(G22B, i2)) / 5.);
  o[17] = (lerp(ay, TXr(G22B, i1), TXr(G22B, i2))) * 1.2;
  return 0;
}
#undef TXr

/* Tabd_yeq0 lerps (reference Tabd.c:118-189): out = v[4], x, xsq, fx, fy, ell1, ell3, dell1dx,
 * dell1dycb, dg1dx, ddg1dxdx, dg2dx, dg1dycb, ddg1dxdycb.  Returns 1 when |v| > XX[last]. */
DEV int yeq0_prims(const DT &T, const double *xv, double *o)
{
  const double xsq = scal(xv, xv);
  const double x = sqrt(fabs(xsq));
  int ia = bsrch(T.XX, x, 0, T.nstpx);
  if (x > T.XX[T.nstpx - 1]) return 1;
  if (ia > T.nstpx - 2) ia = T.nstpx - 2; /* x on the last node: use the last cell */
  const int ib = ia + 1;
  const double *TY = T.TY;
#define TYr(r, i) TY[(r) * NXTAY + (i)]
  const double ax = (T.XX[ib] - x) / (T.XX[ib] - T.XX[ia]);
  const double xa = T.XX[ia], xb = T.XX[ib];
  const double xa2 = xa * xa, xa3 = xa2 * xa, xa4 = xa3 * xa, xb2 = xb * xb, xb3 = xb2 * xb, xb4 = xb3 * xb;
  for (int i = 0; i < 4; i++) o[i] = xv[i];
  o[4] = x; o[5] = xsq;
  o[6] = 0.5 * lerp(ax, TYr(alpha0dx_0p, ia), TYr(alpha0dx_0p, ib));
  o[7] = (lerp(ax, TYr(alpha0_1p, ia), TYr(alpha0_1p, ib)));
  o[8] = 4 * (lerp(ax, TYr(alpha1_0p, ia) / (xa4), TYr(alpha1_0p, ib) / (xb4))) / 3.;
  o[9] = lerp(ax, TYr(beta4_1p, ia) / (xa2), TYr(beta4_1p, ib) / (xb2));
  o[10] = 4 * lerp(ax, (TYr(alpha1dx_0p, ia) - 4. * TYr(alpha1_0p, ia) / xa) / xa4,
                   (TYr(alpha1dx_0p, ib) - 4. * TYr(alpha1_0p, ib) / xb) / xb4) / 3.;
  o[11] = 2 * lerp(ax, (4. * TYr(alpha1_1p, ia) / (3 * xa) - TYr(beta4_1p, ia)) / xa3,
                   (4. * TYr(alpha1_1p, ib) / (3 * xb) - TYr(beta4_1p, ib)) / xb3);
  o[12] = lerp(ax, TYr(alpha3dx_0p, ia), TYr(alpha3dx_0p, ib));
  o[13] = lerp(ax, TYr(alpha3dxdx_0p, ia), TYr(alpha3dxdx_0p, ib));
  o[14] = 2 * lerp(ax, TYr(beta2dx_1p, ia), TYr(beta2dx_1p, ib));
  o[15] = 2 * lerp(ax, TYr(alpha3_1p, ia) - TYr(beta2_1p, ia) / xa, TYr(alpha3_1p, ib) - TYr(beta2_1p, ib) / xb);
  o[16] = 2 * lerp(ax, TYr(beta2_1p, ia) / (xa2) - TYr(beta2dx_1p, ia) / xa + TYr(alpha3dx_1p, ia),
                   TYr(beta2_1p, ib) / (xb2) - TYr(beta2dx_1p, ib) / xb + TYr(alpha3dx_1p, ib));
#undef TYr
  return 0;
}

/* ------------------------------------------------------------------ classification + sort */

/* key 0: special points, out of grid, grid edges; 1 + ix*nstpy + iy: packed path;
 * 1 + nstpx*nstpy + iy: Taylor-x region */
/* Keys of 256 pairs per block, and each pair's slot inside its bucket.  The block sorts its keys in
 * LDS (bitonic, key then position) and takes one atomic per run of equal keys, so that buckets
 * holding a large share of the batch (small |y| cells, the Taylor region) see one atomic per block
 * instead of one per pair.  The counters are G arrays of nb buckets, gcount[g * nb + key], block b
 * counting into array b mod G: concurrent blocks then add into different arrays (measured three
 * times faster than the G counters of a bucket side by side).  Launch with 256 threads. */
__global__ void __launch_bounds__(256)
k_classify(DT T, size_t n, const double *__restrict__ x, int xs, const double *__restrict__ y, int ys, int mode,
           unsigned *__restrict__ key, unsigned *__restrict__ rank, unsigned *__restrict__ gcount, int G, size_t nb)
{
  __shared__ unsigned sk[256], sp[256], sbase[256];
  __shared__ unsigned long long swb[4];
  const int t = threadIdx.x;
  const size_t i = (size_t)blockIdx.x * 256 + t;
  unsigned k = 0xFFFFFFFFu; /* past the end of the batch: sorts last, never counted */
  if (i < n) {
    double xv[4], yv[4];
    for (int c = 0; c < 4; c++) { xv[c] = x[xs * i + c]; yv[c] = y[ys * i + c]; }
    k = 0;
    if (!is_zero(xv) && !is_zero(yv) && !x_is_y(xv, yv)) {
      const Inv I = set_inv(T, xv, yv);
      /* mode KQED_SORT_XNODE keys the packed path by x node only and the Taylor region as one bucket */
      const bool cell = mode == KQED_SORT_CELL;
      if (I.x < T.XX[0]) k = 1 + (unsigned)(T.nstpx * T.nstpy) + (cell ? (unsigned)I.INVy.idx : 0u);
      else if (I.x < T.XX[T.nstpx - 1] && I.y < T.YY[T.nstpy - 1])
        k = 1 + (unsigned)(I.INVx.idx * T.nstpy + (cell ? I.INVy.idx : 0));
    }
    key[i] = k;
  }
  sk[t] = k;
  sp[t] = t;
  __syncthreads();
  for (int size = 2; size <= 256; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const int o = t ^ stride;
      if (o > t) {
        const bool up = (t & size) == 0;
        const unsigned ka = sk[t], kb = sk[o], pa = sp[t], pb = sp[o];
        const bool gt = ka > kb || (ka == kb && pa > pb);
        if (gt == up) { sk[t] = kb; sk[o] = ka; sp[t] = pb; sp[o] = pa; }
      }
      __syncthreads();
    }
  /* runs of equal keys: one atomic per run for the run's length.  The run boundaries of each
   * wavefront are one ballot mask; a run's start is the last boundary at or before t and its end
   * the first boundary after t (in this wavefront's mask or the next ones'). */
  const unsigned kt = sk[t];
  const bool bound = t == 0 || sk[t - 1] != kt;
  const unsigned long long bm = __ballot(bound);
  const int lane = t & 63, wv = t >> 6;
  if (lane == 0) swb[wv] = bm;
  __syncthreads();
  const unsigned long long upto = lane == 63 ? ~0ull : (2ull << lane) - 1; /* bits 0..lane */
  int e = 256;
  if (bm & ~upto) {
    e = wv * 64 + __builtin_ctzll(bm & ~upto);
  } else {
    for (int w2 = wv + 1; w2 < 4; w2++)
      if (swb[w2]) { e = w2 * 64 + __builtin_ctzll(swb[w2]); break; }
  }
  /* bucket kt counts in G sub-counters gcount[g * nb + kt], g = block mod G: a bucket that most
   * blocks touch takes 1/G of the atomics on each address */
  if (kt != 0xFFFFFFFFu && bound) sbase[t] = atomicAdd(&gcount[(size_t)(blockIdx.x % G) * nb + kt], (unsigned)(e - t));
  __syncthreads();
  if (kt != 0xFFFFFFFFu) {
    int s0 = 0;
    if (bm & upto) {
      s0 = wv * 64 + 63 - __builtin_clzll(bm & upto);
    } else {
      for (int w2 = wv - 1; w2 >= 0; w2--)
        if (swb[w2]) { s0 = w2 * 64 + 63 - __builtin_clzll(swb[w2]); break; }
    }
    rank[(size_t)blockIdx.x * 256 + sp[t]] = sbase[s0] + (unsigned)(t - s0);
  }
}

__global__ void k_fill_u32(unsigned *__restrict__ a, size_t n, unsigned v)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

/* Exclusive prefix of the nb bucket counts, in two launches over tiles of 1024 counters: k_scan_tiles
 * turns each tile into its own exclusive prefix and writes the tile's total to tsum[tile];
 * k_scan_add adds the sum of the preceding tiles' totals.  Counter i = G key + g (bucket-major, the
 * order of the starts) is read from the classification's arrays at gcount[g nkey + key]; with G = 1
 * gcount may be count itself.  Coalesced stores. */
__global__ void __launch_bounds__(1024) k_scan_tiles(unsigned *count, unsigned nb, unsigned *__restrict__ tsum,
                                                     const unsigned *gcount, unsigned G, unsigned nkey)
{
  __shared__ unsigned wsum[16];
  const unsigned t = threadIdx.x, lane = t & 63, w = t >> 6;
  const unsigned i = blockIdx.x * 1024 + t;
  const unsigned v = i < nb ? gcount[(size_t)(i % G) * nkey + i / G] : 0u;
  unsigned inc = v; /* inclusive prefix within the wavefront */
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const unsigned u = __shfl_up(inc, off, 64);
    if (lane >= (unsigned)off) inc += u;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  unsigned base = 0;
  for (unsigned k = 0; k < w; k++) base += wsum[k];
  if (i < nb) count[i] = base + inc - v;
  if (t == 1023) {
    unsigned tot = 0;
    for (unsigned k = 0; k < 16; k++) tot += wsum[k];
    tsum[blockIdx.x] = tot;
  }
}

__global__ void __launch_bounds__(1024) k_scan_add(unsigned *__restrict__ count, unsigned nb, const unsigned *__restrict__ tsum)
{
  /* the preceding tiles' totals, summed by the whole block (integer sums: any order is exact) */
  __shared__ unsigned wsum[16];
  const unsigned t = threadIdx.x;
  unsigned part = 0;
  for (unsigned k = t; k < blockIdx.x; k += 1024) part += tsum[k];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
  if ((t & 63) == 0) wsum[t >> 6] = part;
  __syncthreads();
  unsigned off = 0;
  for (int k = 0; k < 16; k++) off += wsum[k];
  const unsigned i = blockIdx.x * 1024 + t;
  if (i < nb && blockIdx.x) count[i] += off;
}

/* start[b] is the exclusive prefix of the bucket counts: pair i goes to start[key] + its rank */
__global__ void k_scatter(size_t n, const unsigned *__restrict__ key, const unsigned *__restrict__ rank,
                          const unsigned *__restrict__ start, unsigned *__restrict__ perm, int G)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  /* the sub-counter of pair i's classification block (256 pairs per block) */
  perm[start[(size_t)key[i] * G + (i >> 8) % G] + rank[i]] = (unsigned)i;
}

/* ------------------------------------------------------------------ interpolation kernel */
/* Evaluation plan of one pair (reference kernels.c:82-231): slot 0 holds L0(x,y) and slots 1-3
 * the subtraction terms of L1, L2, L3 and M-L2, with their coefficients. */
DEV void plan_tabd(const DT &T, int kind, const double *v, double *payload, unsigned &kinds, int slot)
{
  int bad = 0;
  if (kind == KIND_XEQ0 || kind == KIND_XEQ0_M) bad = xeq0_prims(T, v, payload);
  else bad = yeq0_prims(T, v, payload);
  if (bad) kind = KIND_NONE;
  kinds |= (unsigned)kind << (3 * slot);
}

/* L0(a,b) into slot `slot`: regime dispatch of reference kernels.c:82-115 */
template <int CLS>
DEV void plan_L0(const DT &T, const double *a, const double *b, double *rec, unsigned &kinds, int slot,
                 const unsigned char *smask, bool fuse = false, bool *alg_done = nullptr, double (*tw)[17] = nullptr)
{
  const bool az = is_zero(a), bz = is_zero(b);
  double *pl = slot == 0 ? rec + KQED_REC_MAIN : rec + KQED_REC_AUX + KQED_AUX_LEN * (slot - 1);
  if (az && bz) return;
  if (az) { plan_tabd(T, KIND_XEQ0, b, pl, kinds, slot); return; }
  if (bz) { plan_tabd(T, KIND_YEQ0, a, pl, kinds, slot); return; }
  if (x_is_y(a, b)) { plan_tabd(T, KIND_YEQ0_MULAM, a, pl, kinds, slot); return; }
  if (slot != 0) return; /* only the main slot is ever a standard evaluation */
  if (CLS == 1 && smask) {
    /* packed-path pair in a bin where the surrogate passed its error check: only the invariants
     * here, the weight functions come from the network */
    const Inv I = set_inv(T, a, b);
    if (smask[surr_bin(T, I.x, I.y)]) {
      Stage1 *st = (Stage1 *)pl;
      write_header(st, I);
      st->need_alg = 1.0;
      kinds |= (unsigned)KIND_STD | KQED_SURR_BIT;
      return;
    }
  }
  bool done = false;
  if (interp_eval<CLS>(T, a, b, (Stage1 *)pl, fuse, done, tw) == 0) kinds |= (unsigned)KIND_STD;
  if (alg_done) *alg_done = done;
}

DEV void wave_sync();

/* occupancy: the packed-path and Taylor-region kernels run at two waves per SIMD (the packed one
 * spills a little: measured faster than one wave with more registers; the Taylor one measured 6 %
 * faster than at three waves, where it spilled ~450 registers), the generic kernel at one */
template <int CLS> struct InterpWaves { static constexpr int v = 1; };
template <> struct InterpWaves<1> { static constexpr int v = 2; };
template <> struct InterpWaves<2> { static constexpr int v = 2; };

/* Sorted position i (input position when perm is null) in the window [c0, c1) of this pass and in
 * the class segment [lo, hi), lo = bounds[ilo] (0 when ilo < 0), hi = bounds[ihi] (n when ihi < 0).
 * The record of position i is rec[i - c0]; it carries the input index p.  Pair p is
 * (x[xs*p .. +4], y[ys*p .. +4]): a stride of 0 repeats one vector for every pair. */
template <int CLS>
__global__ void __launch_bounds__(KQED_INTERP_THREADS, InterpWaves<CLS>::v)
k_interp(DT T, size_t n, const double *__restrict__ x, int xs, const double *__restrict__ y, int ys,
         const unsigned *__restrict__ perm, const unsigned *__restrict__ bounds, int ilo, int ihi, size_t c0, size_t c1,
         int which, double M, double *__restrict__ rec, unsigned *__restrict__ kinds, const unsigned char *smask, int fuse)
{
  size_t lo = ilo < 0 ? 0 : bounds[ilo], hi = ihi < 0 ? n : bounds[ihi];
  lo = lo > c0 ? lo : c0;
  hi = hi < c1 ? hi : c1;
  const size_t i = lo + (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if constexpr (CLS == 0) {
    if (i >= hi) return;
  }
  /* the packed-path and Taylor-region kernels end with a wavefront-wide copy of the records */
  const bool act = i < hi;
  if (__ballot(act) == 0) return;
  /* CLS 1, 2: the wavefront's LDS tile through which the records' header and the algebra's outputs
   * reach HBM in whole lines (wave_store16) */
  __shared__ double tb[KQED_INTERP_THREADS / 64][64][17];
  double(*tw)[17] = CLS == 0 ? nullptr : tb[threadIdx.x >> 6];
  const size_t ri = i - c0;
  double *r = rec + ri * KQED_REC_LEN;
  double h[16];
  if (act) {
    const size_t p = perm ? perm[i] : i;
    double xv[4], yv[4];
    for (int k = 0; k < 4; k++) { xv[k] = x[xs * p + k]; yv[k] = y[ys * p + k]; }
    unsigned kd = 0;
    const double z[4] = { 0.0, 0.0, 0.0, 0.0 };
    double s1 = 0.0, s2 = 0.0, s3 = 0.0;
    switch (which) {
    case 1: /* L1 = L0(x,y) - L0(x,x)/2 - L0(y,y)/2 */
      plan_L0<CLS>(T, xv, xv, r, kd, 1, nullptr); plan_L0<CLS>(T, yv, yv, r, kd, 2, nullptr); s1 = -0.5; s2 = -0.5; break;
    case 2: /* L2 = L0(x,y) - L0(0,y) - L0(x,0) */
      plan_L0<CLS>(T, z, yv, r, kd, 1, nullptr); plan_L0<CLS>(T, xv, z, r, kd, 2, nullptr); s1 = -1.0; s2 = -1.0; break;
    case 3: /* L3 = L0(x,y) - L0(x,x) + L0(0,x) - L0(0,y) */
      plan_L0<CLS>(T, xv, xv, r, kd, 1, nullptr); plan_L0<CLS>(T, z, xv, r, kd, 2, nullptr);
      plan_L0<CLS>(T, z, yv, r, kd, 3, nullptr);
      s1 = -1.0; s2 = 1.0; s3 = -1.0; break;
    case 4: /* M-L2 = L0(x,y) - e^{-Mx^2/2}(...L0(0,y)) - e^{-My^2/2}(...L0(x,0)) */
      if (!is_zero(yv)) plan_tabd(T, KIND_XEQ0_M, yv, r + KQED_REC_AUX, kd, 1);
      if (!is_zero(xv)) plan_tabd(T, KIND_YEQ0_M, xv, r + KQED_REC_AUX + KQED_AUX_LEN, kd, 2);
      break;
    default: break;
    }
    bool alg_done = false;
    plan_L0<CLS>(T, xv, yv, r, kd, 0, smask, fuse != 0, &alg_done, tw);
    for (int k = 0; k < 4; k++) { h[KQED_REC_X + k] = xv[k]; h[KQED_REC_Y + k] = yv[k]; }
    h[KQED_REC_COEF + 0] = 1.0;
    h[KQED_REC_COEF + 1] = s1;
    h[KQED_REC_COEF + 2] = s2;
    h[KQED_REC_COEF + 3] = s3;
    h[KQED_REC_M] = M;
    /* the M-L2 damping factors (exp(-0) = 1 exactly, so M = 0 skips the exponentials) */
    h[KQED_REC_GX] = M != 0.0 ? exp(-M * (xv[0] * xv[0] + xv[1] * xv[1] + xv[2] * xv[2] + xv[3] * xv[3]) / 2.) : 1.0;
    h[KQED_REC_GY] = M != 0.0 ? exp(-M * (yv[0] * yv[0] + yv[1] * yv[1] + yv[2] * yv[2] + yv[3] * yv[3]) / 2.) : 1.0;
    h[KQED_REC_P] = (double)p;
    if constexpr (CLS == 0)
      for (int k = 0; k < 16; k++) r[k] = h[k];
    kinds[ri] = kd;
    /* fuse: the chain-rule algebra of a standard evaluation right here (no separate pass) */
    if (fuse && !alg_done && (kd & 7u) == (unsigned)KIND_STD && !(kd & KQED_SURR_BIT))
      algebra_main(r + KQED_REC_MAIN, xv, yv);
  }
  /* the whole wavefront stores every active lane's header */
  if constexpr (CLS != 0) wave_store16(tw, r - (threadIdx.x & 63) * KQED_REC_LEN, h, __ballot(act));
}

template __global__ void k_interp<0>(DT, size_t, const double *, int, const double *, int, const unsigned *,
                                     const unsigned *, int, int, size_t, size_t, int, double, double *, unsigned *,
                                     const unsigned char *, int);
template __global__ void k_interp<1>(DT, size_t, const double *, int, const double *, int, const unsigned *,
                                     const unsigned *, int, int, size_t, size_t, int, double, double *, unsigned *,
                                     const unsigned char *, int);
template __global__ void k_interp<2>(DT, size_t, const double *, int, const double *, int, const unsigned *,
                                     const unsigned *, int, int, size_t, size_t, int, double, double *, unsigned *,
                                     const unsigned char *, int);

/* ------------------------------------------------------------------ tensor entries */
/* Per-pair helper vectors of the standard path (shared by many tensor entries), each evaluated with
 * the expression of the reference (chnr_dS.c:86-93, chnr_dT.c:259-263, chnr_dV.c:286-340):
 *   0 xhat, 4 yhat, 8 c2v, 12 c4v, 16 D1, 20 D2, 24 D3 (= D5), 28 D4, 32 D6, 36 D7, 40 D8,
 *   44 astuff[alpha], 48 bstuff[beta], 52/56/60 T_x coefficient vectors, 64/68/72 T_y ones,
 *   76 c1[alpha][beta] of get_dv */
#define NH 92
/* rx = 1/x, ry = 1/y: the helpers multiply by the reciprocals where the reference divides by x
 * and y (the results agree to rounding) */
/* Helper vectors of the standard assembly (per pair, in LDS), groups of four (index i = j & 3):
 *   0 xh, 1 yh, 2 c2, 3 c4, 4..8 x P[26+2k] + y P[27+2k], 9..10 (x P[21+2k] + y P[22+2k])/|x|,
 *   11 (P21+P22) xh + (P23+P24) c4, 12 P21 xh + P25 yh + P23 e12,
 *   13..15 xh P[6+3k] + c4 P[8+3k], 16..18 yh P[7+3k] + c2 P[8+3k],
 * then 16 quadratic ones, j = 76 + 4a + b, which std_entry forms where it uses them (from the
 * unit vectors it reads anyway: no divergent helper branch).  Groups 0..18 are all ((U_i a + V_i b) + W_i c) s with
 * the vectors U, V, W and scalars a, b, c, s read from the record (offsets from its start, HDESC),
 * so the 64 lanes run one instruction stream; the vectors xh, yh, c2, c4, e12 and the sums come
 * from the algebra kernel, and every helper keeps its expression's order of operations. */
#define HR(k) (KQED_REC_MAIN + (k))                  /* primitive k */
#define HV(k) (KQED_REC_MAIN + KQED_HELP_BASE + (k)) /* algebra-kernel value k */
enum { HX = KQED_REC_X, HY = KQED_REC_Y, HXH = HV(0), HYH = HV(4), HC2 = HV(8), HC4 = HV(12), HE12 = HV(16),
       H2122 = HV(20), H2324 = HV(21), HRX = HV(22), HRY = HV(23), HONE = HV(24), HZERO = HV(25) };
/* U, a, V, b, W, c, s */
__constant__ unsigned char HDESC[19][7] = {
  { HXH, HONE, HX, HZERO, HX, HZERO, HONE },       { HYH, HONE, HX, HZERO, HX, HZERO, HONE },
  { HC2, HONE, HX, HZERO, HX, HZERO, HONE },       { HC4, HONE, HX, HZERO, HX, HZERO, HONE },
  { HX, HR(26), HY, HR(27), HX, HZERO, HONE },     { HX, HR(28), HY, HR(29), HX, HZERO, HONE },
  { HX, HR(30), HY, HR(31), HX, HZERO, HONE },     { HX, HR(32), HY, HR(33), HX, HZERO, HONE },
  { HX, HR(34), HY, HR(35), HX, HZERO, HONE },     { HX, HR(21), HY, HR(22), HX, HZERO, HRX },
  { HX, HR(23), HY, HR(24), HX, HZERO, HRX },      { HXH, H2122, HC4, H2324, HX, HZERO, HONE },
  { HXH, HR(21), HYH, HR(25), HE12, HR(23), HONE }, { HXH, HR(6), HC4, HR(8), HX, HZERO, HONE },
  { HXH, HR(9), HC4, HR(11), HX, HZERO, HONE },    { HXH, HR(12), HC4, HR(14), HX, HZERO, HONE },
  { HYH, HR(7), HC2, HR(8), HX, HZERO, HONE },     { HYH, HR(10), HC2, HR(11), HX, HZERO, HONE },
  { HYH, HR(13), HC2, HR(14), HX, HZERO, HONE } };

/* Two pairs' values side by side (the two-pair assembly keeps every LDS slot as one D2): the
 * operators act on both components with the scalar code's operations and order, so a D2
 * evaluation gives each pair exactly its scalar result. */
struct __align__(16) D2 { double x, y; };
struct __align__(8) F2 { float x, y; };
DEV D2 operator+(D2 a, D2 b) { return { a.x + b.x, a.y + b.y }; }
DEV D2 operator-(D2 a, D2 b) { return { a.x - b.x, a.y - b.y }; }
DEV D2 operator*(D2 a, D2 b) { return { a.x * b.x, a.y * b.y }; }
DEV D2 operator+(double a, D2 b) { return { a + b.x, a + b.y }; }
DEV D2 operator+(D2 a, double b) { return { a.x + b, a.y + b }; }
DEV D2 operator-(double a, D2 b) { return { a - b.x, a - b.y }; }
DEV D2 operator-(D2 a, double b) { return { a.x - b, a.y - b }; }
DEV D2 operator*(double a, D2 b) { return { a * b.x, a * b.y }; }
DEV D2 operator*(D2 a, double b) { return { a.x * b, a.y * b }; }
DEV D2 operator/(D2 a, double b) { return { a.x / b, a.y / b }; }
DEV D2 operator+(D2 a) { return a; }
DEV D2 &operator+=(D2 &a, D2 b) { a = a + b; return a; }
DEV D2 fma(double a, D2 b, D2 c) { return { fma(a, b.x, c.x), fma(a, b.y, c.y) }; }
DEV D2 fma(D2 a, D2 b, D2 c) { return { fma(a.x, b.x, c.x), fma(a.y, b.y, c.y) }; }
DEV D2 operator/(D2 a, D2 b) { return { a.x / b.x, a.y / b.y }; }

/* group helper j < 76; d: its descriptor packed as bytes U|a|V|b (d.x) and W|c|s (d.y); R: the
 * record in LDS */
template <typename V>
DEV V std_helper(const V *R, int j, uint2 d)
{
  const int i = j & 3;
  const V u = R[(d.x & 0xFF) + i] * R[(d.x >> 8) & 0xFF];
  const V v = R[((d.x >> 16) & 0xFF) + i] * R[d.x >> 24];
  const V w = R[(d.y & 0xFF) + i] * R[(d.y >> 8) & 0xFF];
  return ((u + v) + w) * R[(d.y >> 16) & 0xFF];
}

/* S/T/V entry c of the standard path: 0-63 Vv, 64-127 Txv, 128-191 Tyv, 192-195 Sxv, 196-199 Syv */
template <typename V>
DEV V std_entry(const V *P, const V *xv, const V *yv, const V *h, int c)
{
  if (c >= 192) {
    const int b = c & 3;
    if (c < 196) return h[b] * P[0] + h[12 + b] * P[2];
    return h[4 + b] * P[1] + h[8 + b] * P[2];
  }
  const int alf = (c >> 4) & 3, bet = (c >> 2) & 3, dta = c & 3;
  const double d_ab = (alf == bet), d_bd = (bet == dta), d_ad = (alf == dta);
  if (c >= 64) {
    const V xsq = P[18], ysq = P[19], xdoty = P[20];
    const V t1 = d_ab * yv[dta] + d_bd * yv[alf] - 0.5 * d_ad * yv[bet];
    const V t2 = d_ab * xv[dta] + d_bd * xv[alf] - 0.5 * d_ad * xv[bet];
    const V t3 = xv[alf] * xv[dta] - d_ad * xsq / 4;
    const V t4 = yv[alf] * yv[dta] - d_ad * ysq / 4;
    const V t5 = xv[alf] * yv[dta] + yv[alf] * xv[dta] - 0.5 * xdoty * d_ad;
    if (c < 128) return +t2 * P[3] + t1 * P[5] + t3 * h[52 + bet] + t4 * h[56 + bet] + t5 * h[60 + bet];
    return +t1 * P[4] + t2 * P[5] + t3 * h[64 + bet] + t4 * h[68 + bet] + t5 * h[72 + bet];
  }
  const V xa = h[alf], xb = h[bet], ya = h[4 + alf], yb = h[4 + bet];
  const V c4a = h[12 + alf], c2c4 = h[8 + bet] + h[12 + bet];
  const V xaxb = xa * xb, xayb = xa * yb;
  V f = d_bd * h[44 + alf];
  f += d_ad * h[48 + bet];
  f += h[16 + dta] * xaxb + h[20 + dta] * xayb;
  f += (h[24 + dta] * xb + h[28 + dta] * yb) * c4a;
  f += (h[24 + dta] * xa + h[32 + dta] * c4a) * c2c4;
  f += h[36 + dta] * (d_ab - xaxb);
  /* the quadratic helper (alf, bet) from the values at hand: the unit vectors x/|x|, y/|y| at alf
   * and bet are h[0..8), so this is the helper's own expression with its own operands */
  const V cb = P[17], rx = P[KQED_HELP_BASE + 22], ry = P[KQED_HELP_BASE + 23];
  const V yayb = ya * yb, yaxb = ya * xb;
  const V quad = ((cb * (3.0 * xaxb - d_ab) - xayb - yaxb) * rx + (d_ab + xayb * cb - xaxb - yayb) * ry);
  return f + h[40 + dta] * quad;
}

/* Tabd_xeq0 tensors (reference Tabd.c:71-111); h = y/|y|.  c: 0-63 tI, 64-127 tII, 128-191 tIII.
 * V: double, or D2 for two pairs side by side */
template <typename V>
DEV V xeq0_entry(const V *P, const V *h, int c)
{
  const V *yv = P;
  const V y = P[4], ysq = P[5], fx = P[6], fy = P[7], ell2 = P[8], dell2adx = P[9], dell2ady = P[10];
  const V ell3a = P[11], dg2dx = P[12], ddg2dxdy = P[13], dg1dx = P[14], dg1dy = P[15], phi1 = P[16], phi2 = P[17];
  const int t = c >> 6, alf = (c >> 4) & 3, bet = (c >> 2) & 3, dta = c & 3;
  const V ha = h[alf], hb = h[bet], hd = h[dta];
  const double dab = alf == bet, dbd = bet == dta, dad = alf == dta;
  const V ysq_4 = ysq / 4.;
  const V E = (ell3a + ell2), Dd2 = (dell2adx + dell2ady), Dd1 = (dg1dy + dg1dx), Dd12 = (dg1dx + dg2dx),
          Dd3 = (y * ddg2dxdy - dg2dx);
  if (t == 2) {
    const V dgSUM = ((hb * fx / 2) + (hb * fy / 2));
    const V tIIIsum = dgSUM - 0.5 * (yv[bet] * E) - ysq_4 * (hb * Dd2);
    return (yv[alf] * yv[dta] * (hb * Dd2) + dab * (yv[dta] * E) + dbd * (yv[alf] * E) + dad * tIIIsum);
  }
  if (t == 1) {
    const V yalf2dx = ha * (dell2adx);
    const V tIIsum = (ha * fx / 2) - 0.5 * (yv[alf] * ell3a) - ysq_4 * yalf2dx;
    return (yv[bet] * yv[dta] * yalf2dx + dbd * (tIIsum) + (dab * (yv[dta] * ell3a) + dad * (yv[bet] * ell3a)));
  }
  const V tIsum = 2 * (dab * phi1 + ha * hb * phi2);
  return (dbd * (ha * Dd12) + (hd * dg2dx) * dab + (hb * Dd3) * ha * hd + (hb * Dd1) * dad + yv[dta] * tIsum);
}

/* Tabd_yeq0 tensors (reference Tabd.c:190-217, loop form); h = x/|x| */
template <typename V>
DEV V yeq0_entry(const V *P, const V *h, int c)
{
  const V *xv = P;
  const V xsq = P[5], fx = P[6], fy = P[7], ell1 = P[8], ell3 = P[9], dell1dx = P[10];
  const V dell1dycb = P[11], dg1dx = P[12], ddg1dxdx = P[13], dg2dx = P[14], dg1dycb = P[15], ddg1dxdycb = P[16];
  const int t = c >> 6, alf = (c >> 4) & 3, bet = (c >> 2) & 3, dta = c & 3;
  const V ha = h[alf], hb = h[bet], hd = h[dta];
  const double dab = alf == bet, dbd = bet == dta, dad = alf == dta;
  const V xsq_4 = xsq / 4., E = ell1 + ell3, Dd1 = dell1dx + dell1dycb, Dd2 = (ddg1dxdx + ddg1dxdycb),
          Dd3 = (dg1dycb - dg2dx), Dd4 = (dg1dx + dg1dycb);
  if (t == 2) {
    const V Dg0SUM = ((hb * fx / 2) + (hb * fy / 2) - 0.5 * (xv[bet] * E));
    const V xad = (xv[alf] * xv[dta] - xsq_4 * dad);
    return ((xad * hb * Dd1 + dab * (xv[dta] * E) + dbd * (xv[alf] * E) + Dg0SUM * dad));
  }
  if (t == 1) {
    const V xhd1dx = ha * (dell1dx);
    const V tIIsum = ((ha * fx / 2) - 0.5 * (xv[alf] * ell1));
    return ((xv[bet] * xv[dta] - xsq_4 * dbd) * xhd1dx + dab * (xv[dta] * ell1) + dad * (xv[bet] * ell1) + dbd * (tIIsum));
  }
  return ((dad * hb + dbd * ha + dab * hd - ha * hb * hd) * Dd4 + ha * (hb * xv[dta] * Dd2 - dbd * Dd3));
}

/* ------------------------------------------------------------------ assembly kernel */
/* One wavefront per pair: the pair's record, helper vectors, tensors and its 384 outputs live in the
 * wavefront's own LDS area; the three assembly maps are shared by the block.  Lane l, slot m
 * computes output row wmRow[q][m][l] of map q; outputs are combined per slot in LDS and stored
 * as 384 contiguous doubles (6 coalesced 512-byte stores per lane set). */
#define WREC KQED_REC_LEN
#define WH 92
#define WHG 76 /* helpers formed in the helper phase; the 16 quadratic ones are formed where used */
#define WT 384

DEV void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* LDS slots of a lane's standard tensor entries c = lane + 64 k (k < 4, c < 200), one byte each */
DEV unsigned std_slots(const DT &T, int lane)
{
  unsigned w = 0;
#pragma unroll
  for (int k = 0; k < 4; k++)
    if (lane + 64 * k < 200) w |= (unsigned)T.wmSlot[lane + 64 * k] << (8 * k);
  return w;
}
DEV int std_slot(unsigned w, int k) { return (w >> (8 * k)) & 0xFF; }

/* One row of a lane: KB entries (byte offset col*16 of the column's two-pair slot | upper 16 bits
 * of the FP64 coefficient << 16; one pair's FP64 or FP32 rows use half or a quarter of it), unrolled so that every LDS read of the row is in flight at once.  The FP64
 * coefficient is the entry with its low half cleared, as the high word of a double. */
template <int KB, typename R>
DEV R wm_dot(const unsigned *__restrict__ e, int lane, const R *Tv)
{
  unsigned w[KB];
#pragma unroll
  for (int k = 0; k < KB; k++) w[k] = e[64 * k + lane];
  R t[KB];
#pragma unroll
  for (int k = 0; k < KB; k++)
    t[k] = *(const R *)((const char *)Tv + ((w[k] & 0xFFFFu) >> (sizeof(R) == 8 ? 1 : 2)));
  R s = 0;
#pragma unroll
  for (int k = 0; k < KB; k++) {
    const double c = __hiloint2double((int)(w[k] & 0xFFFF0000u), 0);
    s = fma((R)c, t[k], s);
  }
  return s;
}


/* wm_dot for two pairs side by side: one entry decode and one 16-byte tensor read per term */
template <int KB>
DEV D2 wm_dot2(const unsigned *__restrict__ e, int lane, const D2 *Tv)
{
  unsigned w[KB];
#pragma unroll
  for (int k = 0; k < KB; k++) w[k] = e[64 * k + lane];
  D2 t[KB];
#pragma unroll
  for (int k = 0; k < KB; k++) t[k] = *(const D2 *)((const char *)Tv + (w[k] & 0xFFFFu));
  D2 s = D2{ 0.0, 0.0 };
#pragma unroll
  for (int k = 0; k < KB; k++) {
    const double c = __hiloint2double((int)(w[k] & 0xFFFF0000u), 0);
    s = fma(c, t[k], s);
  }
  return s;
}

/* wm_dot2 with the entries already in registers */
template <int KB>
DEV D2 wm_dot2r(const unsigned *w, const D2 *Tv)
{
  D2 t[KB];
#pragma unroll
  for (int k = 0; k < KB; k++) t[k] = *(const D2 *)((const char *)Tv + (w[k] & 0xFFFFu));
  D2 s = D2{ 0.0, 0.0 };
#pragma unroll
  for (int k = 0; k < KB; k++) {
    const double c = __hiloint2double((int)(w[k] & 0xFFFF0000u), 0);
    s = fma(c, t[k], s);
  }
  return s;
}

/* the same in FP32 (the map of the FP32 layouts): 8-byte elements, so the byte offset is col * 8,
 * half the entry's */
template <int KB>
DEV D2 wm_dot2(const unsigned *__restrict__ e, int lane, const F2 *Tv)
{
  unsigned w[KB];
#pragma unroll
  for (int k = 0; k < KB; k++) w[k] = e[64 * k + lane];
  F2 t[KB];
#pragma unroll
  for (int k = 0; k < KB; k++) t[k] = *(const F2 *)((const char *)Tv + ((w[k] & 0xFFFFu) >> 1));
  float sx = 0.0f, sy = 0.0f;
#pragma unroll
  for (int k = 0; k < KB; k++) {
    const float c = (float)__hiloint2double((int)(w[k] & 0xFFFF0000u), 0);
    sx = fmaf(c, t[k].x, sx);
    sy = fmaf(c, t[k].y, sy);
  }
  return D2{ (double)sx, (double)sy };
}

/* FP64 sum over the 64 lanes of a wavefront (every lane active), result in every lane */
DEV double wave_sum(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

/* R is the arithmetic type of the sparse assembly maps (the 200 tensor entries and their linear
 * combinations); the slot combination of L1..L3 / M-L2 and every output sum stay FP64.
 * Output modes: layout bit 0 selects out[384][ntot] (else out[q][384]), bit 1 FP32 outputs (else
 * FP64); layout 16 (contraction) writes the FP64 dot product out[q] = sum_c L_c Tc[q][c] with the
 * FP32 site tensor Tc, layout 64 the same with an FP64 tensor (Tc then points to doubles).  With Wc
 * (layout 16 and a single standard L0 slot) the dot takes the linear
 * shortcut sum_c L_c T_c = sum_j u_j W_j, u the 200 tensor entries and W = A^T T per site. */
template <typename MT>
__global__ void __launch_bounds__(64 * KQED_ASM_WAVES)
k_assemble(DT T, size_t n, const double *__restrict__ rec, const unsigned *__restrict__ kinds, void *__restrict__ outv,
           int layout, int nslots, size_t ntot, const void *__restrict__ Tc, const double *__restrict__ Wc,
           const unsigned *__restrict__ split, size_t c0)
{
  typedef MT R_t;
  /* split: only the records before sorted position split[0] (the special pairs; the two-pair
   * kernel takes the rest) */
  if (split) {
    const size_t sp = split[0] > c0 ? split[0] - c0 : 0;
    n = sp < n ? sp : n;
  }
  if ((size_t)blockIdx.x * KQED_ASM_WAVES >= n) return;
  __shared__ unsigned sE[KQED_WM_MAX];
  __shared__ unsigned short sRow[3 * 384];
  __shared__ double sW[KQED_ASM_WAVES][WREC + WH + 201 + WT];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < T.wmTotal; i += blockDim.x) sE[i] = T.wmE[i];
  for (int i = threadIdx.x; i < 3 * 384; i += blockDim.x) sRow[i] = T.wmRow[i];
  __syncthreads();
  double *R = sW[w], *H = R + WREC, *O = H + WH + 201;
  R_t *Tv = (R_t *)(H + WH);
  const size_t stride = (size_t)gridDim.x * KQED_ASM_WAVES;
  /* the next pair's record and kinds are loaded into registers while the current pair is assembled,
   * so the wavefront never waits on a record load from HBM */
  static_assert(WREC <= 192, "record prefetch covers three doubles per lane");
  static_assert(WH == 92 && KQED_HELP_BASE + 26 <= 64, "helper layout");
  /* descriptors of this lane's two helpers (j = lane, lane + 64) */
  auto hdesc = [](int g) {
    const unsigned char *e = HDESC[g < 19 ? g : 0];
    return make_uint2(e[0] | e[1] << 8 | e[2] << 16 | (unsigned)e[3] << 24, e[4] | e[5] << 8 | e[6] << 16);
  };
  const uint2 hd0 = hdesc(lane >> 2), hd1 = hdesc((lane + 64) >> 2);
  const unsigned slw = std_slots(T, lane); /* LDS slots of this lane's standard entries */
  size_t p = (size_t)blockIdx.x * KQED_ASM_WAVES + w;
  /* one evaluation slot reads the header and main payload only (80 of the record's doubles) */
  const int used = nslots == 1 ? KQED_REC_AUX : WREC;
  double nx0 = 0.0, nx1 = 0.0, nx2 = 0.0;
  unsigned nk = 0;
  if (p < n) {
    const double *rp = rec + p * WREC;
    nx0 = rp[lane];
    if (lane + 64 < used) nx1 = rp[lane + 64];
    if (lane + 128 < used) nx2 = rp[lane + 128];
    nk = kinds[p];
  }
  /* The outputs of a pair wait in O and go to HBM at the start of the next iteration, after the
   * next record is staged: vector loads and stores share one completion counter, so a store issued
   * just before that staging would make the wavefront wait for its write to complete. */
  size_t pq = 0;        /* pair whose outputs are pending in O (dot layouts: in pdot) */
  double pdot = 0.0;
  bool pending = false;
  auto flush = [&]() {
    if (layout == 16 || layout == 64) {
      if (lane == 0) ((double *)outv)[pq] = pdot;
    } else if (layout & 2) {
      float *out = (float *)outv;
      if (layout & 1) for (int i = lane; i < 384; i += 64) __builtin_nontemporal_store((float)O[i], out + (size_t)i * ntot + pq);
      else for (int i = lane; i < 384; i += 64) __builtin_nontemporal_store((float)O[i], out + pq * 384 + i);
    } else {
      double *out = (double *)outv;
      if (layout & 1) for (int i = lane; i < 384; i += 64) __builtin_nontemporal_store(O[i], out + (size_t)i * ntot + pq);
      else for (int i = lane; i < 384; i += 64) __builtin_nontemporal_store(O[i], out + pq * 384 + i);
    }
  };
  for (; p < n; p += stride) {
    R[lane] = nx0;
    if (lane + 64 < used) R[lane + 64] = nx1;
    if (lane + 128 < used) R[lane + 128] = nx2;
    const unsigned kall = __builtin_amdgcn_readfirstlane(nk);
    if (pending) flush();
    if (p + stride < n) {
      const double *rp = rec + (p + stride) * WREC;
      nx0 = __builtin_nontemporal_load(rp + lane);
      if (lane + 64 < used) nx1 = __builtin_nontemporal_load(rp + lane + 64);
      if (lane + 128 < used) nx2 = __builtin_nontemporal_load(rp + lane + 128);
      nk = kinds[p + stride];
    }
    if (lane == 0) Tv[200] = 0.0;
    /* one evaluation slot (L0): each lane's six rows go from registers straight into O (or the
     * dot); several slots accumulate in O, which starts at zero */
    const bool direct = nslots == 1;
    const bool dotl = layout == 16 || layout == 64;
    if (!direct)
      for (int i = lane; i < 384; i += 64) O[i] = 0.0;
    wave_sync();
    const size_t q = (size_t)R[KQED_REC_P];
    double dot = 0.0;
    bool have_dot = false;
    /* rows v[m] of map mq of the single slot: into O, or straight into the dot */
    auto emit = [&](const double *v, int mq) {
      if (dotl) {
        double tq[6];
#pragma unroll
        for (int m = 0; m < 6; m++) {
          const int r = sRow[mq * 384 + 64 * m + lane];
          tq[m] = layout == 16 ? (double)((const float *)Tc)[q * 384 + r] : ((const double *)Tc)[q * 384 + r];
        }
        double part = 0.0;
#pragma unroll
        for (int m = 0; m < 6; m++) part = fma(v[m], tq[m], part);
        dot = wave_sum(part);
        have_dot = true;
        return;
      }
#pragma unroll
      for (int m = 0; m < 6; m++) O[sRow[mq * 384 + 64 * m + lane]] = v[m];
    };
    for (int slot = 0; slot < nslots; slot++) {
      const unsigned kd = (kall >> (3 * slot)) & 7u;
      if (kd == KIND_NONE) {
        if (direct) {
          const double z[6] = { 0, 0, 0, 0, 0, 0 };
          emit(z, 0);
        }
        continue;
      }
      const double *pl = slot == 0 ? R + KQED_REC_MAIN : R + KQED_REC_AUX + KQED_AUX_LEN * (slot - 1);
      if (kd == KIND_STD) {
        H[lane] = std_helper(R, lane, hd0);
        if (lane + 64 < WHG) H[lane + 64] = std_helper(R, lane + 64, hd1);
      } else if (lane < 4) {
        H[lane] = pl[lane] / pl[4];
      }
      wave_sync();
      if (kd == KIND_STD) {
        for (int k = 0, c = lane; c < 200; k++, c += 64)
          Tv[std_slot(slw, k)] = std_entry(pl, R + KQED_REC_X, R + KQED_REC_Y, H, c);
      } else if (kd == KIND_XEQ0 || kd == KIND_XEQ0_M) {
        for (int c = lane; c < 192; c += 64) Tv[c] = xeq0_entry(pl, H, c);
      } else {
        for (int c = lane; c < 192; c += 64) Tv[c] = yeq0_entry(pl, H, c);
      }
      wave_sync();
      if (Wc && kd == KIND_STD) {
        /* contraction shortcut: only slot 0 exists (nslots == 1) */
        const double *wq = Wc + q * KQED_WCOLS;
        double part = 0.0;
        for (int k = 0, c = lane; c < 200; k++, c += 64) {
          const int cc = kqed_wcol(c);
          part = fma((double)Tv[std_slot(slw, k)], cc >= 0 ? wq[cc] : 0.0, part);
        }
        dot = wave_sum(part);
        have_dot = true;
        continue;
      }
      const int mq = kd == KIND_STD ? 0 : kd == KIND_YEQ0_MULAM ? 2 : 1;
      double v[6];
      {
        /* slot m of map mq starts at 64 * (41 mq + 13 + 9 + ...) entries (checked on the host) */
        const unsigned *e = sE + mq * 64 * KQED_WM_PER_LANE;
        v[0] = wm_dot<13>(e, lane, Tv);
        v[1] = wm_dot<9>(e + 64 * 13, lane, Tv);
        v[2] = wm_dot<6>(e + 64 * 22, lane, Tv);
        v[3] = wm_dot<6>(e + 64 * 28, lane, Tv);
        v[4] = wm_dot<4>(e + 64 * 34, lane, Tv);
        v[5] = wm_dot<3>(e + 64 * 38, lane, Tv);
      }
      if (kd == KIND_XEQ0_M || kd == KIND_YEQ0_M) {
        /* reference kernels.c:292-323: row r = (rs, mu, nu, la) of the slot takes the alpha sum
         * S = sum_a v_a K(rs, a, nu, la) (x = 0 term; K(rs, mu, a, la) for the y = 0 term) over other
         * rows of the same L0(0,y) / L0(x,0) kernel.  The lanes hold those rows, so the 96 sums are
         * gathered in LDS (in the tensor region, free once the rows are computed) in four rounds, one
         * per a: every sum gets one term per round, in the order a = 0..3. */
        const bool xm = kd == KIND_XEQ0_M;
        const double Mv = R[KQED_REC_M];
        const double g = xm ? R[KQED_REC_GX] : R[KQED_REC_GY];
        const double *vv = xm ? R + KQED_REC_X : R + KQED_REC_Y;
        double *S = (double *)Tv;
        wave_sync(); /* every lane is done reading the tensors */
        for (int i = lane; i < 96; i += 64) S[i] = 0.0;
        wave_sync();
        for (int a = 0; a < 4; a++) {
#pragma unroll
          for (int m = 0; m < 6; m++) {
            const int r = sRow[mq * 384 + 64 * m + lane];
            const int rs = r >> 6, mu = (r >> 4) & 3, nu = (r >> 2) & 3, la = r & 3;
            if ((xm ? mu : nu) == a) {
              double *sp = S + 16 * rs + 4 * (xm ? nu : mu) + la;
              *sp = *sp + vv[a] * v[m];
            }
          }
          wave_sync();
        }
#pragma unroll
        for (int m = 0; m < 6; m++) {
          const int r = sRow[mq * 384 + 64 * m + lane];
          const int rs = r >> 6, mu = (r >> 4) & 3, nu = (r >> 2) & 3, la = r & 3;
          const double sum = S[16 * rs + 4 * (xm ? nu : mu) + la];
          O[r] -= g * (v[m] - Mv * vv[xm ? mu : nu] * sum);
        }
      } else if (direct) {
        emit(v, mq);
      } else {
        const double cf = R[KQED_REC_COEF + slot];
#pragma unroll
        for (int m = 0; m < 6; m++) {
          const int r = sRow[mq * 384 + 64 * m + lane];
          O[r] = slot == 0 ? v[m] : fma(v[m], cf, O[r]); /* atomic_kernel_Saxpy, kernels.c:52-78 */
        }
      }
      wave_sync();
    }
    if (dotl && !have_dot) {
      double part = 0.0;
      if (layout == 16) {
        const float *tq = (const float *)Tc + q * 384;
        for (int i = lane; i < 384; i += 64) part = fma(O[i], (double)tq[i], part);
      } else {
        const double *tq = (const double *)Tc + q * 384;
        for (int i = lane; i < 384; i += 64) part = fma(O[i], tq[i], part);
      }
      dot = wave_sum(part);
    }
    pq = q;
    pdot = dot;
    pending = true;
    wave_sync();
  }
  if (pending) flush();
}

/* ------------------------------------------------------------------ contraction helpers */
/* Pulled-back site tensors of the contraction shortcut: W[i][wcol(c)] = sum over column c of the
 * standard map of coefficient * T_i[row], entries in the column's (row) order, for the 152 columns a
 * row reads.  One wavefront per site, the sites of a wavefront strided by the grid; the site's 384
 * FP32 values are staged in LDS (the next site's are already loaded).  The columns are dealt to the
 * lanes in three slots of fixed length (KQED_PB_K, longest columns first, padding entries with
 * coefficient 0), the lane's 48 entries held in registers: every step of a slot is one LDS read and
 * one fma in all lanes, with no per-column loop.  With colpart, the block also sums its sites'
 * tensors per column in a fixed order (each wavefront in site order, then the four wavefronts in
 * order) into colpart[block][384]. */
#ifndef KQED_PB_WAVES
#define KQED_PB_WAVES 4 /* waves per SIMD of the pull-back (its 48 entries stay in registers) */
#endif
__global__ void __launch_bounds__(256, KQED_PB_WAVES)
k_contract_weights(size_t nx, const float *__restrict__ Tc, const unsigned *__restrict__ ent,
                   const int *__restrict__ colx, double *__restrict__ W, double *__restrict__ colpart)
{
  __shared__ float st[4][384];
  __shared__ double psum[4][384];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  constexpr int K0 = 26, K1 = 18, K2 = 4;
  static_assert(K0 + K1 + K2 == KQED_PB_TOTAL, "pull-back slot lengths");
  unsigned e[KQED_PB_TOTAL];
#pragma unroll
  for (int k = 0; k < KQED_PB_TOTAL; k++) e[k] = ent[64 * k + lane];
  const int cx0 = colx[lane], cx1 = colx[64 + lane], cx2 = colx[128 + lane];
  double cs[6] = { 0, 0, 0, 0, 0, 0 };
  const size_t stride = (size_t)gridDim.x * 4;
  size_t i = (size_t)blockIdx.x * 4 + w;
  float nt[6];
  if (i < nx)
#pragma unroll
    for (int k = 0; k < 6; k++) nt[k] = Tc[i * 384 + lane + 64 * k];
  auto step = [&](unsigned u, double s) {
    return fma(__hiloint2double((int)(u & 0xFFFF0000u), 0), (double)st[w][u & 0xFFFFu], s);
  };
  for (; i < nx; i += stride) {
#pragma unroll
    for (int k = 0; k < 6; k++) {
      st[w][lane + 64 * k] = nt[k];
      cs[k] += (double)nt[k];
    }
    if (i + stride < nx)
#pragma unroll
      for (int k = 0; k < 6; k++) nt[k] = Tc[(i + stride) * 384 + lane + 64 * k];
    wave_sync();
    /* the entries are decoded at each use: decoded once outside the loop they would take three
     * registers each instead of one */
#pragma unroll
    for (int k = 0; k < KQED_PB_TOTAL; k++) asm volatile("" : "+v"(e[k]));
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int k = 0; k < K0; k++) s0 = step(e[k], s0);
#pragma unroll
    for (int k = 0; k < K1; k++) s1 = step(e[K0 + k], s1);
#pragma unroll
    for (int k = 0; k < K2; k++) s2 = step(e[K0 + K1 + k], s2);
    double *o = W + i * KQED_WCOLS;
    o[cx0] = s0;
    o[cx1] = s1;
    if (cx2 >= 0) o[cx2] = s2;
    wave_sync();
  }
  if (!colpart) return;
#pragma unroll
  for (int k = 0; k < 6; k++) psum[w][lane + 64 * k] = cs[k];
  __syncthreads();
  for (int c = threadIdx.x; c < 384; c += blockDim.x)
    colpart[(size_t)blockIdx.x * 384 + c] = ((psum[0][c] + psum[1][c]) + psum[2][c]) + psum[3][c];
}

/* Deterministic sum of in[0, n): block b reduces the contiguous range [b*per, (b+1)*per) in a fixed
 * order into out[b] (launch with 256 threads). */
__global__ void k_block_sum(size_t n, size_t per, const double *__restrict__ in, double *__restrict__ out)
{
  __shared__ double part[256];
  const size_t lo = (size_t)blockIdx.x * per, hi = lo + per < n ? lo + per : n;
  double s = 0.0;
  for (size_t i = lo + threadIdx.x; i < hi; i += 256) s += in[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) part[threadIdx.x] += part[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = part[0];
}

/* sum[c] = sum over the nb partials part[b][c] in a fixed order: block c, thread t adds the partials
 * b = t, t + 256, ..., then a fixed tree over the 256 threads */
__global__ void __launch_bounds__(256) k_colsum_final(int nb, const double *__restrict__ part, double *__restrict__ sum)
{
  __shared__ double r[256];
  const int c = blockIdx.x, t = threadIdx.x;
  double s = 0.0;
  for (int b = t; b < nb; b += 256) s += part[(size_t)b * 384 + c];
  r[t] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) r[t] += r[t + h];
    __syncthreads();
  }
  if (t == 0) sum[c] = r[0];
}

/* s[j] = sigma[j] + ca * (La[j] . sumT) + cc0 * consts[0] + cc1 * consts[1]  (one wavefront per j) */
__global__ void k_combine(int ny, const double *__restrict__ sigma, const double *__restrict__ La, double ca,
                          const double *__restrict__ sumT, const double *__restrict__ consts, double cc0, double cc1,
                          double *__restrict__ s)
{
  const int j = blockIdx.x, lane = threadIdx.x;
  if (j >= ny) return;
  double part = 0.0;
  if (La)
    for (int c = lane; c < 384; c += 64) part = fma(La[(size_t)j * 384 + c], sumT[c], part);
  const double dot = wave_sum(part);
  if (lane == 0) {
    double v = sigma[j];
    if (La) v = fma(ca, dot, v);
    if (consts) { v = fma(cc0, consts[0], v); v = fma(cc1, consts[1], v); }
    s[j] = v;
  }
}

template __global__ void k_assemble<double>(DT, size_t, const double *, const unsigned *, void *, int, int, size_t,
                                            const void *, const double *, const unsigned *, size_t);
template __global__ void k_assemble<float>(DT, size_t, const double *, const unsigned *, void *, int, int, size_t,
                                           const void *, const double *, const unsigned *, size_t);

/* Two pairs per wavefront: the standard-path assembly of one evaluation slot (L0) with FP64 or
 * FP32 output (MT), for the records from sorted position split[0] on (the packed path and the Taylor-x
 * region, whose kinds are only KIND_STD or KIND_NONE).  Records, helpers, tensor entries and
 * outputs of the two pairs lie side by side in LDS (D2), so every LDS access, map-entry decode and
 * index computation serves both pairs; each pair's arithmetic is the one-pair kernel's, in the
 * same order, so the outputs are bit-identical to it.  The output area aliases the helper and
 * tensor areas (they are dead once the rows are formed).  layout bit 0: out[384][ntot] (else
 * out[q][384]); the element type is MT. */
template <int AW, typename MT, bool WDOT>
__global__ void __launch_bounds__(64 * AW)
k_assemble2(DT T, size_t n, const unsigned *__restrict__ split, size_t c0, const double *__restrict__ rec,
            const unsigned *__restrict__ kinds, void *__restrict__ outv, int layout, size_t ntot,
            const double *__restrict__ Wc)
{
  __shared__ unsigned sE[64 * KQED_WM_PER_LANE]; /* the standard map (map 0 comes first) */
  __shared__ unsigned short sRow[384];
  __shared__ D2 sW[AW][KQED_REC_AUX + WT];
  size_t lo = split[0] > c0 ? split[0] - c0 : 0;
  lo = lo < n ? lo : n;
  const size_t np = (n - lo + 1) / 2; /* record pairs */
  if ((size_t)blockIdx.x * AW >= np) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if constexpr (!WDOT) {
    /* the dot of the contraction shortcut never applies the map */
    for (int i = threadIdx.x; i < 64 * KQED_WM_PER_LANE; i += blockDim.x) sE[i] = T.wmE[i];
    for (int i = threadIdx.x; i < 384; i += blockDim.x) sRow[i] = T.wmRow[i];
    __syncthreads();
  }
  /* MT: arithmetic of the sparse map (the FP32 layouts run it in FP32, as the one-pair kernel) */
  typedef typename std::conditional<sizeof(MT) == 8, D2, F2>::type TV;
  D2 *R = sW[w], *H = R + KQED_REC_AUX, *O = H;
  TV *Tv = (TV *)(H + WH);
  const size_t stride = (size_t)gridDim.x * AW;
  auto hdesc = [](int g) {
    const unsigned char *e = HDESC[g < 19 ? g : 0];
    return make_uint2(e[0] | e[1] << 8 | e[2] << 16 | (unsigned)e[3] << 24, e[4] | e[5] << 8 | e[6] << 16);
  };
  const uint2 hd0 = hdesc(lane >> 2), hd1 = hdesc((lane + 64) >> 2);
  const unsigned slw = std_slots(T, lane); /* LDS slots of this lane's standard entries */
  /* compact weight indices of this lane's entries c = lane + 64 k, one byte each (0xFF: a column no
   * row reads, weight 0) */
  unsigned wcx = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int cc = lane + 64 * k < 200 ? kqed_wcol(lane + 64 * k) : -1;
    wcx |= (cc < 0 ? 0xFFu : (unsigned)cc) << (8 * k);
  }
  /* the first row slot's 13 map entries stay in registers (a longer set spills) */
  unsigned er[13];
#pragma unroll
  for (int k = 0; k < 13; k++) er[k] = WDOT ? 0u : sE[64 * k + lane];
  /* the next two records (header and main payload: 80 doubles) wait in registers */
  double a0 = 0.0, a1 = 0.0, b0 = 0.0, b1 = 0.0;
  unsigned ka = KIND_NONE, kb = KIND_NONE;
  auto load = [&](size_t qq) {
    const size_t ia = lo + 2 * qq, ib = ia + 1;
    const double *ra = rec + ia * WREC;
    a0 = __builtin_nontemporal_load(ra + lane);
    a1 = lane < KQED_REC_AUX - 64 ? __builtin_nontemporal_load(ra + lane + 64) : 0.0;
    ka = kinds[ia];
    if (ib < n) {
      const double *rb = ra + WREC;
      b0 = __builtin_nontemporal_load(rb + lane);
      b1 = lane < KQED_REC_AUX - 64 ? __builtin_nontemporal_load(rb + lane + 64) : 0.0;
      kb = kinds[ib];
    } else {
      b0 = b1 = 0.0;
      kb = KIND_NONE;
    }
  };
  size_t q = (size_t)blockIdx.x * AW + w;
  if (q < np) load(q);
  /* outputs wait in O until the next records are staged (loads and stores share one counter); the
   * outputs and records are touched once, so they are stored and loaded non-temporally (they would
   * only push the interpolation's table nodes out of the caches) */
  size_t pa = 0, pb = 0;
  double pda = 0.0, pdb = 0.0;
  bool pending = false, pbv = false;
  auto flush = [&]() {
    if constexpr (WDOT) {
      if (lane == 0) {
        ((double *)outv)[pa] = pda;
        if (pbv) ((double *)outv)[pb] = pdb;
      }
      return;
    }
    if constexpr (sizeof(MT) == 8) {
      if (!(layout & 1)) {
        /* AoS FP64: two consecutive outputs of a pair per lane, one 16-byte store each */
        typedef double dbl2 __attribute__((ext_vector_type(2)));
        double *out = (double *)outv;
        for (int i = 2 * lane; i < 384; i += 128) {
          const D2 o0 = O[i], o1 = O[i + 1];
          const dbl2 va = { o0.x, o1.x };
          __builtin_nontemporal_store(va, (dbl2 *)(out + pa * 384 + i));
          if (pbv) {
            const dbl2 vb = { o0.y, o1.y };
            __builtin_nontemporal_store(vb, (dbl2 *)(out + pb * 384 + i));
          }
        }
        return;
      }
    }
    for (int i = lane; i < 384; i += 64) {
      const D2 o = O[i];
      MT *out = (MT *)outv;
      if (layout & 1) {
        __builtin_nontemporal_store((MT)o.x, out + (size_t)i * ntot + pa);
        if (pbv) __builtin_nontemporal_store((MT)o.y, out + (size_t)i * ntot + pb);
      } else {
        __builtin_nontemporal_store((MT)o.x, out + pa * 384 + i);
        if (pbv) __builtin_nontemporal_store((MT)o.y, out + pb * 384 + i);
      }
    }
  };
  for (; q < np; q += stride) {
    R[lane] = D2{ a0, b0 };
    if (lane < KQED_REC_AUX - 64) R[lane + 64] = D2{ a1, b1 };
    const unsigned kda = __builtin_amdgcn_readfirstlane(ka) & 7u, kdb = __builtin_amdgcn_readfirstlane(kb) & 7u;
    const bool hasb = lo + 2 * q + 1 < n;
    if (pending) flush();
    if (q + stride < np) load(q + stride);
    wave_sync();
    if (lane == 0) Tv[200] = TV{ 0, 0 };
    /* the pairs' output indices are wave-uniform (scalar registers: the weight loads below then
     * take one address register each) */
    const size_t qa = __builtin_amdgcn_readfirstlane((unsigned)R[KQED_REC_P].x),
                 qb = __builtin_amdgcn_readfirstlane((unsigned)R[KQED_REC_P].y);
    /* the contraction shortcut's site weights: loaded now, so that the helper and tensor phases hide
     * their latency */
    double wpa[4], wpb[4];
    if constexpr (WDOT) {
      const double *wa = Wc + qa * KQED_WCOLS, *wb = Wc + qb * KQED_WCOLS;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const unsigned cc = (wcx >> (8 * k)) & 0xFFu;
        wpa[k] = cc != 0xFFu ? wa[cc] : 0.0;
        wpb[k] = cc != 0xFFu ? wb[cc] : 0.0;
      }
    }
    H[lane] = std_helper(R, lane, hd0);
    if (lane + 64 < WHG) H[lane + 64] = std_helper(R, lane + 64, hd1);
    wave_sync();
    for (int k = 0, c = lane; c < 200; k++, c += 64) {
      const D2 t = std_entry(R + KQED_REC_MAIN, R + KQED_REC_X, R + KQED_REC_Y, H, c);
      Tv[std_slot(slw, k)] = TV{ (MT)t.x, (MT)t.y };
    }
    wave_sync();
    if constexpr (WDOT) {
      /* contraction shortcut (layout 16, FP64 map type): dot of the 200 tensor entries with the
       * site's pulled-back weights, as the one-pair kernel forms it; an empty slot gives 0 */
      double sa = 0.0, sb = 0.0;
      for (int k = 0, c = lane; c < 200; k++, c += 64) {
        const TV t = Tv[std_slot(slw, k)];
        sa = fma((double)t.x, wpa[k], sa);
        sb = fma((double)t.y, wpb[k], sb);
      }
      sa = wave_sum(sa);
      sb = wave_sum(sb);
      pda = kda == KIND_STD ? sa : 0.0;
      pdb = kdb == KIND_STD ? sb : 0.0;
      pa = qa;
      pb = qb;
      pbv = hasb;
      pending = true;
      wave_sync();
      continue;
    }
    D2 v[6];
    {
      const unsigned *e = sE;
      if constexpr (sizeof(MT) == 8) v[0] = wm_dot2r<13>(er, (const D2 *)Tv);
      else v[0] = wm_dot2<13>(e, lane, Tv);
      v[1] = wm_dot2<9>(e + 64 * 13, lane, Tv);
      v[2] = wm_dot2<6>(e + 64 * 22, lane, Tv);
      v[3] = wm_dot2<6>(e + 64 * 28, lane, Tv);
      v[4] = wm_dot2<4>(e + 64 * 34, lane, Tv);
      v[5] = wm_dot2<3>(e + 64 * 38, lane, Tv);
    }
    wave_sync(); /* every lane is done with the helper and tensor areas, which O overlays */
#pragma unroll
    for (int m = 0; m < 6; m++) {
      D2 o = v[m];
      if (kda != KIND_STD) o.x = 0.0;
      if (kdb != KIND_STD) o.y = 0.0;
      O[sRow[64 * m + lane]] = o;
    }
    pa = qa;
    pb = qb;
    pbv = hasb;
    pending = true;
    wave_sync();
  }
  if (pending) flush();
}

template __global__ void k_assemble2<KQED_ASM2_WAVES, double, false>(DT, size_t, const unsigned *, size_t, const double *,
                                                                     const unsigned *, void *, int, size_t, const double *);
template __global__ void k_assemble2<KQED_ASM2_WAVES, float, false>(DT, size_t, const unsigned *, size_t, const double *,
                                                                    const unsigned *, void *, int, size_t, const double *);
template __global__ void k_assemble2<KQED_ASM2D_WAVES, double, true>(DT, size_t, const unsigned *, size_t, const double *,
                                                                    const unsigned *, void *, int, size_t, const double *);

/* The Tabd slots 1..nslots-1 of two records (kinds packed 3 bits per slot) can be applied map by
 * map on combined inputs: every slot is empty or a plain Tabd kind (no M-L2 term), and where both
 * records have a slot set they agree on its kind. */
DEV bool tabd_mergeable(unsigned ka, unsigned kb, int nslots)
{
  for (int s = 1; s < nslots; s++) {
    const unsigned a = (ka >> (3 * s)) & 7u, b = (kb >> (3 * s)) & 7u;
    const bool pa = a == KIND_NONE || a == KIND_XEQ0 || a == KIND_YEQ0 || a == KIND_YEQ0_MULAM;
    const bool pb = b == KIND_NONE || b == KIND_XEQ0 || b == KIND_YEQ0 || b == KIND_YEQ0_MULAM;
    if (!pa || !pb || (a != KIND_NONE && b != KIND_NONE && a != b)) return false;
  }
  return true;
}

/* Several evaluation slots (L1, L2, L3, M-L2) in sorted order past the special pairs, two records per
 * wavefront as in k_assemble2: a slot is evaluated once for both records when their kinds agree
 * (past the special pairs they do, unless a slot of one record is KIND_NONE, whose component is then
 * left unchanged), and once per record when they differ.  Each record's arithmetic is the one-pair
 * kernel's, in the same order (slot 0 stored, later slots fma'd with their coefficient, the dot
 * layouts summed over O in lane order, the M-L2 slots' alpha sums added in the order a = 0..3), so
 * the outputs are bit-identical to it.
 * Layouts: 0/1 out[q][384] / out[384][ntot] of MT, 16 / 64 the contraction dot with an FP32 /
 * FP64 site tensor Tc (MT double). */
template <int AW, typename MT>
__global__ void __launch_bounds__(64 * AW)
k_assemble2m(DT T, size_t n, const unsigned *__restrict__ split, size_t c0, const double *__restrict__ rec,
             const unsigned *__restrict__ kinds, void *__restrict__ outv, int layout, int nslots, size_t ntot,
             const void *__restrict__ Tc)
{
  __shared__ unsigned sE[KQED_WM_MAX];
  __shared__ unsigned short sRow[3 * 384];
  __shared__ D2 sW[AW][WREC + WH + 201 + WT];
  size_t lo = split[0] > c0 ? split[0] - c0 : 0;
  lo = lo < n ? lo : n;
  const size_t np = (n - lo + 1) / 2; /* record pairs */
  if ((size_t)blockIdx.x * AW >= np) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < T.wmTotal; i += blockDim.x) sE[i] = T.wmE[i];
  for (int i = threadIdx.x; i < 3 * 384; i += blockDim.x) sRow[i] = T.wmRow[i];
  __syncthreads();
  typedef typename std::conditional<sizeof(MT) == 8, D2, F2>::type TV;
  D2 *R = sW[w], *H = R + WREC, *O = H + WH + 201;
  TV *Tv = (TV *)(H + WH);
  const size_t stride = (size_t)gridDim.x * AW;
  const bool dotl = layout == 16 || layout == 64;
  auto hdesc = [](int g) {
    const unsigned char *e = HDESC[g < 19 ? g : 0];
    return make_uint2(e[0] | e[1] << 8 | e[2] << 16 | (unsigned)e[3] << 24, e[4] | e[5] << 8 | e[6] << 16);
  };
  const uint2 hd0 = hdesc(lane >> 2), hd1 = hdesc((lane + 64) >> 2);
  const unsigned slw = std_slots(T, lane); /* LDS slots of this lane's standard entries */
  unsigned er[13]; /* map 0's first row slot, kept in registers */
#pragma unroll
  for (int k = 0; k < 13; k++) er[k] = sE[64 * k + lane];
  static_assert(WREC > 128 && WREC <= 192, "record staging covers three doubles per lane");
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, b0 = 0.0, b1 = 0.0, b2 = 0.0;
  unsigned ka = KIND_NONE, kb = KIND_NONE;
  auto load = [&](size_t qq) {
    const size_t ia = lo + 2 * qq, ib = ia + 1;
    const double *ra = rec + ia * WREC;
    a0 = __builtin_nontemporal_load(ra + lane);
    a1 = __builtin_nontemporal_load(ra + lane + 64);
    a2 = lane + 128 < WREC ? __builtin_nontemporal_load(ra + lane + 128) : 0.0;
    ka = kinds[ia];
    if (ib < n) {
      const double *rb = ra + WREC;
      b0 = __builtin_nontemporal_load(rb + lane);
      b1 = __builtin_nontemporal_load(rb + lane + 64);
      b2 = lane + 128 < WREC ? __builtin_nontemporal_load(rb + lane + 128) : 0.0;
      kb = kinds[ib];
    } else {
      b0 = b1 = b2 = 0.0;
      kb = KIND_NONE;
    }
  };
  size_t q = (size_t)blockIdx.x * AW + w;
  if (q < np) load(q);
  size_t pa = 0, pb = 0;
  double pda = 0.0, pdb = 0.0;
  bool pending = false, pbv = false;
  auto flush = [&]() {
    if (dotl) {
      if (lane == 0) {
        ((double *)outv)[pa] = pda;
        if (pbv) ((double *)outv)[pb] = pdb;
      }
      return;
    }
    MT *out = (MT *)outv;
    for (int i = lane; i < 384; i += 64) {
      const D2 o = O[i];
      if (layout & 1) {
        __builtin_nontemporal_store((MT)o.x, out + (size_t)i * ntot + pa);
        if (pbv) __builtin_nontemporal_store((MT)o.y, out + (size_t)i * ntot + pb);
      } else {
        __builtin_nontemporal_store((MT)o.x, out + pa * 384 + i);
        if (pbv) __builtin_nontemporal_store((MT)o.y, out + pb * 384 + i);
      }
    }
  };
  for (; q < np; q += stride) {
    R[lane] = D2{ a0, b0 };
    R[lane + 64] = D2{ a1, b1 };
    if (lane + 128 < WREC) R[lane + 128] = D2{ a2, b2 };
    const unsigned kda = __builtin_amdgcn_readfirstlane(ka), kdb = __builtin_amdgcn_readfirstlane(kb);
    const bool hasb = lo + 2 * q + 1 < n;
    if (pending) flush();
    if (q + stride < np) load(q + stride);
    for (int i = lane; i < 384; i += 64) O[i] = D2{ 0.0, 0.0 };
    wave_sync();
    const size_t qa = (size_t)R[KQED_REC_P].x, qb = hasb ? (size_t)R[KQED_REC_P].y : qa;
    for (int slot = 0; slot < nslots; slot++) {
      if (slot == 1 && tabd_mergeable(kda, kdb, nslots)) {
        /* L1, L2, L3: the Tabd maps are linear, so the slots that share a map are combined on their
         * inputs (coefficient times tensor entry, FP64, in slot order; an empty slot adds nothing)
         * and each map is applied once: L1 and L2 need one Tabd map application instead of two, L3
         * two instead of three (reference kernels.c:119-177) */
        for (int mq = 1; mq <= 2; mq++) {
          int sl[3];
          unsigned kk[3];
          int ns = 0;
          for (int s2 = 1; s2 < nslots; s2++) {
            const unsigned a = (kda >> (3 * s2)) & 7u, b = (kdb >> (3 * s2)) & 7u;
            const unsigned k = a != KIND_NONE ? a : b;
            if (k == KIND_NONE || (k == KIND_YEQ0_MULAM ? 2 : 1) != mq) continue;
            sl[ns] = s2;
            kk[ns] = k;
            ns++;
          }
          if (ns == 0) continue;
          /* unit vectors of the slots' arguments, H[4 j + i] */
          if (lane < 4 * ns) {
            const D2 *pl = R + KQED_REC_AUX + KQED_AUX_LEN * (sl[lane >> 2] - 1);
            H[lane] = pl[lane & 3] / pl[4];
          }
          wave_sync();
          D2 acc0 = D2{ 0.0, 0.0 }, acc1 = D2{ 0.0, 0.0 }, acc2 = D2{ 0.0, 0.0 };
          for (int j = 0; j < ns; j++) {
            const int s2 = sl[j];
            const D2 *pl = R + KQED_REC_AUX + KQED_AUX_LEN * (s2 - 1);
            const bool va = ((kda >> (3 * s2)) & 7u) != KIND_NONE, vb = ((kdb >> (3 * s2)) & 7u) != KIND_NONE;
            D2 t0, t1, t2;
            if (kk[j] == KIND_XEQ0) {
              t0 = xeq0_entry(pl, H + 4 * j, lane); t1 = xeq0_entry(pl, H + 4 * j, lane + 64);
              t2 = xeq0_entry(pl, H + 4 * j, lane + 128);
            } else {
              t0 = yeq0_entry(pl, H + 4 * j, lane); t1 = yeq0_entry(pl, H + 4 * j, lane + 64);
              t2 = yeq0_entry(pl, H + 4 * j, lane + 128);
            }
            const D2 cf = R[KQED_REC_COEF + s2];
            /* an empty slot's payload is not read: its component is zero */
            const D2 c = D2{ va ? cf.x : 0.0, vb ? cf.y : 0.0 };
            if (!va) t0.x = t1.x = t2.x = 0.0;
            if (!vb) t0.y = t1.y = t2.y = 0.0;
            acc0 = fma(c, t0, acc0);
            acc1 = fma(c, t1, acc1);
            acc2 = fma(c, t2, acc2);
          }
          Tv[lane] = TV{ (MT)acc0.x, (MT)acc0.y };
          Tv[lane + 64] = TV{ (MT)acc1.x, (MT)acc1.y };
          Tv[lane + 128] = TV{ (MT)acc2.x, (MT)acc2.y };
          if (lane == 0) Tv[200] = TV{ 0, 0 };
          wave_sync();
          const unsigned *e = sE + mq * 64 * KQED_WM_PER_LANE;
          D2 v[6];
          v[0] = wm_dot2<13>(e, lane, Tv);
          v[1] = wm_dot2<9>(e + 64 * 13, lane, Tv);
          v[2] = wm_dot2<6>(e + 64 * 22, lane, Tv);
          v[3] = wm_dot2<6>(e + 64 * 28, lane, Tv);
          v[4] = wm_dot2<4>(e + 64 * 34, lane, Tv);
          v[5] = wm_dot2<3>(e + 64 * 38, lane, Tv);
#pragma unroll
          for (int m = 0; m < 6; m++) {
            const int r = sRow[mq * 384 + 64 * m + lane];
            O[r] = O[r] + v[m];
          }
          wave_sync();
        }
        break;
      }
      const unsigned sa = (kda >> (3 * slot)) & 7u, sb = (kdb >> (3 * slot)) & 7u;
      for (int pass = 0; pass < 2; pass++) {
        /* pass 0: the kind of record a (of b when a's slot is empty); pass 1: b's differing kind */
        unsigned kd;
        if (pass == 0) kd = sa != KIND_NONE ? sa : sb;
        else if (sa != KIND_NONE && sb != KIND_NONE && sa != sb) kd = sb;
        else break;
        if (kd == KIND_NONE) break;
        const bool ua = pass == 0 && sa == kd, ub = sb == kd;
        const D2 *pl = slot == 0 ? R + KQED_REC_MAIN : R + KQED_REC_AUX + KQED_AUX_LEN * (slot - 1);
        /* the maps' zero column, set for every slot: an M-L2 slot's sums overwrite it in the FP32
         * layouts, whose tensor entries are 8 bytes */
        if (lane == 0) Tv[200] = TV{ 0, 0 };
        if (kd == KIND_STD) {
          H[lane] = std_helper(R, lane, hd0);
          if (lane + 64 < WHG) H[lane + 64] = std_helper(R, lane + 64, hd1);
        } else if (lane < 4) {
          H[lane] = pl[lane] / pl[4];
        }
        wave_sync();
        /* Tabd tensors: a lane's three entries c = lane + 64 k are formed before any is stored (no
         * LDS store lies between their reads of the record, so the per-pair values are read and
         * formed once) */
        if (kd == KIND_STD) {
          for (int k = 0, c = lane; c < 200; k++, c += 64) {
            const D2 t = std_entry(pl, R + KQED_REC_X, R + KQED_REC_Y, H, c);
            Tv[std_slot(slw, k)] = TV{ (MT)t.x, (MT)t.y };
          }
        } else if (kd == KIND_XEQ0 || kd == KIND_XEQ0_M) {
          const D2 t0 = xeq0_entry(pl, H, lane), t1 = xeq0_entry(pl, H, lane + 64), t2 = xeq0_entry(pl, H, lane + 128);
          Tv[lane] = TV{ (MT)t0.x, (MT)t0.y };
          Tv[lane + 64] = TV{ (MT)t1.x, (MT)t1.y };
          Tv[lane + 128] = TV{ (MT)t2.x, (MT)t2.y };
        } else {
          const D2 t0 = yeq0_entry(pl, H, lane), t1 = yeq0_entry(pl, H, lane + 64), t2 = yeq0_entry(pl, H, lane + 128);
          Tv[lane] = TV{ (MT)t0.x, (MT)t0.y };
          Tv[lane + 64] = TV{ (MT)t1.x, (MT)t1.y };
          Tv[lane + 128] = TV{ (MT)t2.x, (MT)t2.y };
        }
        wave_sync();
        const int mq = kd == KIND_STD ? 0 : kd == KIND_YEQ0_MULAM ? 2 : 1;
        const unsigned *e = sE + mq * 64 * KQED_WM_PER_LANE;
        D2 v[6];
        if constexpr (sizeof(MT) == 8) {
          if (mq == 0) v[0] = wm_dot2r<13>(er, (const D2 *)Tv);
          else v[0] = wm_dot2<13>(e, lane, Tv);
        } else {
          v[0] = wm_dot2<13>(e, lane, Tv);
        }
        v[1] = wm_dot2<9>(e + 64 * 13, lane, Tv);
        v[2] = wm_dot2<6>(e + 64 * 22, lane, Tv);
        v[3] = wm_dot2<6>(e + 64 * 28, lane, Tv);
        v[4] = wm_dot2<4>(e + 64 * 34, lane, Tv);
        v[5] = wm_dot2<3>(e + 64 * 38, lane, Tv);
        if (kd == KIND_XEQ0_M || kd == KIND_YEQ0_M) {
          /* the M-L2 term (reference kernels.c:292-323).  Sum (rs, nu, la) (x = 0 term; (rs, mu, la)
           * for y = 0) is ((0 + p_0) + p_1) + p_2) + p_3 with p_a = v_a(row) vv[a], as the one-pair
           * kernel's four rounds form it: each row stores its p_a once (a < 3: into buffer a, 96
           * sums each, in the free helper and tensor areas), then the row with a = 3 adds the three
           * and its own p_3 and leaves the sum in buffer 0.  Both records use the same map, so each
           * lane's rows and sums are the same for both. */
          const bool xm = kd == KIND_XEQ0_M;
          const D2 Mv = R[KQED_REC_M];
          const D2 g = xm ? R[KQED_REC_GX] : R[KQED_REC_GY];
          const D2 *vv = xm ? R + KQED_REC_X : R + KQED_REC_Y;
          static_assert(3 * 96 <= WH + 201, "three sum buffers fit the helper and tensor areas");
          D2 *B = H;
          int ia[6], ix[6];
#pragma unroll
          for (int m = 0; m < 6; m++) {
            const int r = sRow[mq * 384 + 64 * m + lane];
            const int rs = r >> 6, mu = (r >> 4) & 3, nu = (r >> 2) & 3, la = r & 3;
            ia[m] = xm ? mu : nu;
            ix[m] = 16 * rs + 4 * (xm ? nu : mu) + la;
          }
          wave_sync(); /* every lane is done reading the helpers and tensors */
#pragma unroll
          for (int m = 0; m < 6; m++)
            if (ia[m] < 3) B[96 * ia[m] + ix[m]] = vv[ia[m]] * v[m];
          wave_sync();
#pragma unroll
          for (int m = 0; m < 6; m++)
            if (ia[m] == 3) B[ix[m]] = ((0.0 + B[ix[m]]) + B[96 + ix[m]]) + B[192 + ix[m]] + vv[3] * v[m];
          wave_sync();
#pragma unroll
          for (int m = 0; m < 6; m++) {
            const int r = sRow[mq * 384 + 64 * m + lane];
            const D2 sum = B[ix[m]];
            const D2 t = O[r] - g * (v[m] - Mv * vv[ia[m]] * sum);
            D2 o = O[r];
            if (ua) o.x = t.x;
            if (ub) o.y = t.y;
            O[r] = o;
          }
        } else {
          const D2 cf = R[KQED_REC_COEF + slot];
#pragma unroll
          for (int m = 0; m < 6; m++) {
            const int r = sRow[mq * 384 + 64 * m + lane];
            D2 o = O[r];
            if (ua) o.x = slot == 0 ? v[m].x : fma(v[m].x, cf.x, o.x); /* atomic_kernel_Saxpy, kernels.c:52-78 */
            if (ub) o.y = slot == 0 ? v[m].y : fma(v[m].y, cf.y, o.y);
            O[r] = o;
          }
        }
        wave_sync();
      }
    }
    double da = 0.0, db = 0.0;
    if (dotl) {
      double sa = 0.0, sb = 0.0;
      if (layout == 16) {
        const float *ta = (const float *)Tc + qa * 384, *tb = (const float *)Tc + qb * 384;
        for (int i = lane; i < 384; i += 64) {
          const D2 o = O[i];
          sa = fma(o.x, (double)ta[i], sa);
          sb = fma(o.y, (double)tb[i], sb);
        }
      } else {
        const double *ta = (const double *)Tc + qa * 384, *tb = (const double *)Tc + qb * 384;
        for (int i = lane; i < 384; i += 64) {
          const D2 o = O[i];
          sa = fma(o.x, ta[i], sa);
          sb = fma(o.y, tb[i], sb);
        }
      }
      da = wave_sum(sa);
      db = wave_sum(sb);
    }
    pa = qa;
    pb = qb;
    pda = da;
    pdb = db;
    pbv = hasb;
    pending = true;
    wave_sync();
  }
  if (pending) flush();
}

template __global__ void k_assemble2m<KQED_ASM2M_WAVES, double>(DT, size_t, const unsigned *, size_t, const double *,
                                                                const unsigned *, void *, int, int, size_t, const void *);
template __global__ void k_assemble2m<KQED_ASM2M_WAVES, float>(DT, size_t, const unsigned *, size_t, const double *,
                                                               const unsigned *, void *, int, int, size_t, const void *);

/* Tabd tensor entry c of a slot's payload pl (unit vector h) for kind k: Tabd_xeq0, Tabd_yeq0, or,
 * for L(v, v) = -L_{mu<->lambda}(v, 0), the Tabd_yeq0 entry that the Tabd map needs at c for the
 * same outputs: the trace reversal Tr{Gam g_l g_a g_n g_b g_m} = -Tr{Gam g_m g_b g_n g_a g_l} turns
 * the mu<->lambda exchange of the rows into alpha<->beta on the inputs with T^II and T^III exchanged,
 * and the sign with it: (T^I, T^II, T^III)_abd <- (T^I, T^III, T^II)_bad. */
DEV D2 tabd_value(unsigned k, const D2 *pl, const D2 *h, int c)
{
  if (k == KIND_XEQ0) return xeq0_entry(pl, h, c);
  if (k == KIND_YEQ0) return yeq0_entry(pl, h, c);
  const int t = c >> 6, swp = 16 * ((c >> 2) & 3) + 4 * ((c >> 4) & 3) + (c & 3);
  return yeq0_entry(pl, h, (t == 0 ? 0 : t == 1 ? 128 : 64) + swp);
}

/* L1, L2, L3 in sorted order past the special pairs, two records per wavefront, with ONE application
 * of the Tabd map (also the one-slot dot passes whose records are all special points: slot 0 of a
 * Tabd kind enters as a Tabd slot with coefficient 1, which gives the one-pair kernel's XEQ0/YEQ0
 * sums exactly).  Both maps of kqed_assembly.c are the Dirac-trace contraction
 * L = sum_A G^A T^(A); the standard map is it on the S/T/V entries folded into
 *   T^I = Vv, T^II_abd = Txv[b][a][d] + 1/4 d_bd Sxv[a], T^III_abd = Txv + Tyv + 1/4 d_ad (Sxv + Syv)[b],
 * and L(v, v) is it on permuted Tabd_yeq0 entries (tabd_value).  The kernel combinations are linear,
 * so (reference kernels.c:119-177)
 *   L1 = M (F(u) - 1/2 s(t(x,0)) - 1/2 s(t(y,0))),  L2 = M (F(u) - t(0,y) - t(x,0)),
 *   L3 = M (F(u) - s(t(x,0)) + t(0,x) - t(0,y)),
 * one gather of 2016 map entries per pair where the slot-by-slot form runs 2280 + 2016 (L1, L2) or
 * 2280 + 2 x 2016 (L3).  The inputs are combined in FP64 (standard part, then the slots in order
 * with their coefficients; an empty slot adds nothing), so the result rounds differently from the
 * one-pair kernel, within the FP64 bar.  Only the rep columns of the folded T^II (b <= d) and
 * T^III (a <= d) are formed: the map reads no other.
 * Layouts as k_assemble2m (0/1 AoS/SoA of MT; 16/64 the contraction dot).  The work area holds the
 * helpers and the standard entries (FP64, natural order), then the combined inputs (MT), then the
 * outputs, each phase overlaying the previous one after a wavefront barrier. */
template <int AW, typename MT>
__global__ void __launch_bounds__(64 * AW)
k_assemble2t(DT T, size_t n, const unsigned *__restrict__ split, size_t c0, const double *__restrict__ rec,
             const unsigned *__restrict__ kinds, void *__restrict__ outv, int layout, int nslots, size_t ntot,
             const void *__restrict__ Tc)
{
  __shared__ unsigned sE[64 * KQED_WM_PER_LANE]; /* the Tabd map (map 1) */
  __shared__ unsigned short sRow[384];
  __shared__ D2 sW[AW][WREC + WT];
  static_assert(WH + 201 <= WT && WHG + 12 <= WH, "helpers, unit vectors and standard entries fit the work area");
  size_t lo = split[0] > c0 ? split[0] - c0 : 0;
  lo = lo < n ? lo : n;
  const size_t np = (n - lo + 1) / 2; /* record pairs */
  if ((size_t)blockIdx.x * AW >= np) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < 64 * KQED_WM_PER_LANE; i += blockDim.x) sE[i] = T.wmE[T.wmOff[1][0] + i];
  for (int i = threadIdx.x; i < 384; i += blockDim.x) sRow[i] = T.wmRow[384 + i];
  __syncthreads();
  typedef typename std::conditional<sizeof(MT) == 8, D2, F2>::type TV;
  D2 *R = sW[w], *H = R + WREC, *Ts = H + WH, *O = H;
  TV *U = (TV *)H;
  const size_t stride = (size_t)gridDim.x * AW;
  const bool dotl = layout == 16 || layout == 64;
  auto hdesc = [](int g) {
    const unsigned char *e = HDESC[g < 19 ? g : 0];
    return make_uint2(e[0] | e[1] << 8 | e[2] << 16 | (unsigned)e[3] << 24, e[4] | e[5] << 8 | e[6] << 16);
  };
  const uint2 hd0 = hdesc(lane >> 2), hd1 = hdesc((lane + 64) >> 2);
  /* the rep columns this lane forms: T^I always, T^II when b <= d, T^III when a <= d */
  const int la = lane >> 4, lb = (lane >> 2) & 3, ld = lane & 3;
  const bool rep2 = lb <= ld, rep3 = la <= ld;
  static_assert(WREC > 128 && WREC <= 192, "record staging covers three doubles per lane");
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, b0 = 0.0, b1 = 0.0, b2 = 0.0;
  unsigned ka = KIND_NONE, kb = KIND_NONE;
  const int used = KQED_REC_AUX + KQED_AUX_LEN * (nslots - 1);
  auto load = [&](size_t qq) {
    const size_t ia = lo + 2 * qq, ib = ia + 1;
    const double *ra = rec + ia * WREC;
    a0 = __builtin_nontemporal_load(ra + lane);
    a1 = __builtin_nontemporal_load(ra + lane + 64);
    a2 = lane + 128 < used ? __builtin_nontemporal_load(ra + lane + 128) : 0.0;
    ka = kinds[ia];
    if (ib < n) {
      const double *rb = ra + WREC;
      b0 = __builtin_nontemporal_load(rb + lane);
      b1 = __builtin_nontemporal_load(rb + lane + 64);
      b2 = lane + 128 < used ? __builtin_nontemporal_load(rb + lane + 128) : 0.0;
      kb = kinds[ib];
    } else {
      b0 = b1 = b2 = 0.0;
      kb = KIND_NONE;
    }
  };
  size_t q = (size_t)blockIdx.x * AW + w;
  if (q < np) load(q);
  size_t pa = 0, pb = 0;
  double pda = 0.0, pdb = 0.0;
  bool pending = false, pbv = false;
  auto flush = [&]() {
    if (dotl) {
      if (lane == 0) {
        ((double *)outv)[pa] = pda;
        if (pbv) ((double *)outv)[pb] = pdb;
      }
      return;
    }
    if constexpr (sizeof(MT) == 8) {
      if (!(layout & 1)) {
        /* AoS FP64: two consecutive outputs of a pair per lane, one 16-byte store each */
        typedef double dbl2 __attribute__((ext_vector_type(2)));
        double *out = (double *)outv;
        for (int i = 2 * lane; i < 384; i += 128) {
          const D2 o0 = O[i], o1 = O[i + 1];
          const dbl2 va = { o0.x, o1.x };
          __builtin_nontemporal_store(va, (dbl2 *)(out + pa * 384 + i));
          if (pbv) {
            const dbl2 vb = { o0.y, o1.y };
            __builtin_nontemporal_store(vb, (dbl2 *)(out + pb * 384 + i));
          }
        }
        return;
      }
    }
    MT *out = (MT *)outv;
    for (int i = lane; i < 384; i += 64) {
      const D2 o = O[i];
      if (layout & 1) {
        __builtin_nontemporal_store((MT)o.x, out + (size_t)i * ntot + pa);
        if (pbv) __builtin_nontemporal_store((MT)o.y, out + (size_t)i * ntot + pb);
      } else {
        __builtin_nontemporal_store((MT)o.x, out + pa * 384 + i);
        if (pbv) __builtin_nontemporal_store((MT)o.y, out + pb * 384 + i);
      }
    }
  };
  for (; q < np; q += stride) {
    R[lane] = D2{ a0, b0 };
    R[lane + 64] = D2{ a1, b1 };
    if (lane + 128 < used) R[lane + 128] = D2{ a2, b2 };
    const unsigned kda = __builtin_amdgcn_readfirstlane(ka), kdb = __builtin_amdgcn_readfirstlane(kb);
    const bool hasb = lo + 2 * q + 1 < n;
    if (pending) flush();
    if (q + stride < np) load(q + stride);
    wave_sync();
    const size_t qa = (size_t)R[KQED_REC_P].x, qb = hasb ? (size_t)R[KQED_REC_P].y : qa;
    const bool sa = (kda & 7u) == KIND_STD, sb = (kdb & 7u) == KIND_STD;
    /* the standard part: helpers, then the 200 S/T/V entries (FP64, column order) */
    D2 u0 = D2{ 0.0, 0.0 }, u1 = u0, u2 = u0;
    if (sa || sb) {
      H[lane] = std_helper(R, lane, hd0);
      if (lane + 64 < WHG) H[lane + 64] = std_helper(R, lane + 64, hd1);
    }
    /* the unit vectors of the Tabd slots' arguments, in the free helper slots H[WHG + 4 (s - 1) + i];
     * a Tabd slot 0 (the one-slot passes of the contraction's x-only sums) in H[WHG + 12 + i] */
    const unsigned k0a = kda & 7u, k0b = kdb & 7u;
    const bool t0a = !sa && k0a != KIND_NONE, t0b = !sb && k0b != KIND_NONE;
    if (lane < 4 * (nslots - 1)) {
      const D2 *pl = R + KQED_REC_AUX + KQED_AUX_LEN * (lane >> 2);
      H[WHG + lane] = pl[lane & 3] / pl[4];
    }
    if ((t0a || t0b) && lane >= 12 && lane < 16) H[WHG + lane] = R[KQED_REC_MAIN + lane - 12] / R[KQED_REC_MAIN + 4];
    wave_sync();
    if (sa || sb) {
      for (int c = lane; c < 200; c += 64) Ts[c] = std_entry(R + KQED_REC_MAIN, R + KQED_REC_X, R + KQED_REC_Y, H, c);
      wave_sync();
      /* fold onto T^I, T^II, T^III (reps only) */
      u0 = Ts[lane];
      if (rep2) {
        u1 = Ts[64 + 16 * lb + 4 * la + ld];
        if (lb == ld) u1 = u1 + 0.25 * Ts[192 + la];
      }
      if (rep3) {
        u2 = Ts[64 + lane] + Ts[128 + lane];
        if (la == ld) u2 = u2 + 0.25 * (Ts[192 + lb] + Ts[196 + lb]);
      }
      if (!sa) { u0.x = 0.0; u1.x = 0.0; u2.x = 0.0; }
      if (!sb) { u0.y = 0.0; u1.y = 0.0; u2.y = 0.0; }
    }
    for (int s2 = t0a || t0b ? 0 : 1; s2 < nslots; s2++) {
      /* slot 0 counts here only as a Tabd kind (payload in the main slot, coefficient 1) */
      const unsigned ks_a = s2 ? (kda >> (3 * s2)) & 7u : t0a ? k0a : KIND_NONE;
      const unsigned ks_b = s2 ? (kdb >> (3 * s2)) & 7u : t0b ? k0b : KIND_NONE;
      if (ks_a == KIND_NONE && ks_b == KIND_NONE) continue;
      const D2 *pl = s2 ? R + KQED_REC_AUX + KQED_AUX_LEN * (s2 - 1) : R + KQED_REC_MAIN;
      const D2 *h = H + WHG + (s2 ? 4 * (s2 - 1) : 12);
      const D2 cf = s2 ? R[KQED_REC_COEF + s2] : D2{ 1.0, 1.0 };
      /* an empty slot's payload is not read: its component is zero */
      const D2 cc = D2{ ks_a != KIND_NONE ? cf.x : 0.0, ks_b != KIND_NONE ? cf.y : 0.0 };
      const unsigned k1 = ks_a != KIND_NONE ? ks_a : ks_b;
      D2 t0 = tabd_value(k1, pl, h, lane);
      D2 t1 = rep2 ? tabd_value(k1, pl, h, lane + 64) : D2{ 0.0, 0.0 };
      D2 t2 = rep3 ? tabd_value(k1, pl, h, lane + 128) : D2{ 0.0, 0.0 };
      if (ks_b != KIND_NONE && ks_b != k1) {
        /* the records' kinds differ in this slot: b's component from its own kind */
        t0.y = tabd_value(ks_b, pl, h, lane).y;
        if (rep2) t1.y = tabd_value(ks_b, pl, h, lane + 64).y;
        if (rep3) t2.y = tabd_value(ks_b, pl, h, lane + 128).y;
      }
      if (ks_a == KIND_NONE) t0.x = t1.x = t2.x = 0.0;
      if (ks_b == KIND_NONE) t0.y = t1.y = t2.y = 0.0;
      u0 = fma(cc, t0, u0);
      u1 = fma(cc, t1, u1);
      u2 = fma(cc, t2, u2);
    }
    wave_sync(); /* every lane is done with the helpers and entries, which U overlays */
    U[lane] = TV{ (MT)u0.x, (MT)u0.y };
    U[lane + 64] = TV{ (MT)u1.x, (MT)u1.y };
    U[lane + 128] = TV{ (MT)u2.x, (MT)u2.y };
    if (lane == 0) U[200] = TV{ 0, 0 };
    wave_sync();
    D2 v[6];
    v[0] = wm_dot2<13>(sE, lane, U);
    v[1] = wm_dot2<9>(sE + 64 * 13, lane, U);
    v[2] = wm_dot2<6>(sE + 64 * 22, lane, U);
    v[3] = wm_dot2<6>(sE + 64 * 28, lane, U);
    v[4] = wm_dot2<4>(sE + 64 * 34, lane, U);
    v[5] = wm_dot2<3>(sE + 64 * 38, lane, U);
    wave_sync(); /* every lane is done with U, which O overlays */
#pragma unroll
    for (int m = 0; m < 6; m++) O[sRow[64 * m + lane]] = v[m];
    double da = 0.0, db = 0.0;
    if (dotl) {
      wave_sync();
      double xa = 0.0, xb = 0.0;
      if (layout == 16) {
        const float *ta = (const float *)Tc + qa * 384, *tb = (const float *)Tc + qb * 384;
        for (int i = lane; i < 384; i += 64) {
          const D2 o = O[i];
          xa = fma(o.x, (double)ta[i], xa);
          xb = fma(o.y, (double)tb[i], xb);
        }
      } else {
        const double *ta = (const double *)Tc + qa * 384, *tb = (const double *)Tc + qb * 384;
        for (int i = lane; i < 384; i += 64) {
          const D2 o = O[i];
          xa = fma(o.x, ta[i], xa);
          xb = fma(o.y, tb[i], xb);
        }
      }
      da = wave_sum(xa);
      db = wave_sum(xb);
    }
    pa = qa;
    pb = qb;
    pda = da;
    pdb = db;
    pbv = hasb;
    pending = true;
    wave_sync();
  }
  if (pending) flush();
}

template __global__ void k_assemble2t<KQED_ASM2T_WAVES, double>(DT, size_t, const unsigned *, size_t, const double *,
                                                                const unsigned *, void *, int, int, size_t, const void *);
template __global__ void k_assemble2t<KQED_ASM2T_WAVES, float>(DT, size_t, const unsigned *, size_t, const double *,
                                                               const unsigned *, void *, int, int, size_t, const void *);

/* ragged groups: ypair[i] = y[g] for off[g] <= i < off[g+1] (binary search over the ng+1 offsets) */
__global__ void k_ragged_expand(size_t n, int ng, const unsigned long long *__restrict__ off,
                                const double *__restrict__ y, double *__restrict__ ypair)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int lo = 0, hi = ng; /* off[lo] <= i < off[hi] */
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= i) lo = mid; else hi = mid;
  }
  for (int k = 0; k < 4; k++) ypair[4 * i + k] = y[4 * lo + k];
}

/* Training data of the surrogate: for every packed-path standard pair of a pass (records before the
 * chain-rule algebra), out[p] = {x, y, cb (post-swap invariants), flag2, 1, the 34 interpolated
 * weight functions}; other pairs get out[p][4] = 0. */
__global__ void k_extract_scalars(DT T, size_t n, const double *__restrict__ rec, const unsigned *__restrict__ kinds,
                                  double *__restrict__ out)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *r = rec + i * KQED_REC_LEN;
  const double *m = r + KQED_REC_MAIN;
  const size_t p = (size_t)r[KQED_REC_P];
  double *o = out + p * KQED_SURR_ROW;
  const Stage1 *st = (const Stage1 *)m;
  const bool packed = (kinds[i] & 7u) == (unsigned)KIND_STD && st->need_alg != 0.0 && st->x >= T.XX[0] &&
                      st->x < T.XX[T.nstpx - 1] && st->y < T.YY[T.nstpy - 1];
  o[0] = st->x; o[1] = st->y; o[2] = st->cb; o[3] = st->flag2; o[4] = packed ? 1.0 : 0.0;
  const int slots[KQED_NSURR] = KQED_SURR_SLOTS;
  for (int k = 0; k < KQED_NSURR; k++) o[5 + k] = packed ? m[slots[k]] : 0.0;
  if (packed && st->flag2 == 0.0) { o[5 + 32] = 0.0; o[5 + 33] = 0.0; }
}

/* bin of the error check for each extracted row (-1 unless a packed-path pair) */
__global__ void k_surr_rowbin(DT T, size_t n, const double *__restrict__ data, int *__restrict__ bin)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *row = data + i * KQED_SURR_ROW;
  bin[i] = row[4] != 0.0 ? surr_bin(T, row[0], row[1]) : -1;
}
