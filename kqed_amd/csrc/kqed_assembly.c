/*
 * kqed_assembly.c -- the 384-output assembly maps of the QED kernel, built from Dirac traces.
 *
 * The reference hand-unrolls two fixed linear maps: CONSTRUCT_FULL_KERNEL (QED_kernel.c:16-423,
 * inputs Vv, Txv, Tyv, Sxv, Syv) and CONSTRUCT_KERNEL_FULL_EQZERO (QED_kernel_xy0.c:8-399,
 * inputs T^I, T^II, T^III).  Both are the contraction
 *   L_[rho sigma];mu nu lambda = sum_{alpha beta delta} G^A T^(A)_{alpha beta delta},  A = I, II, III
 * with Euclidean traces, Gam_{delta rho sigma} = 1/2 {g_delta, [g_rho, g_sigma]}:
 *   G^I   =  1/8 Tr{Gam g_mu g_alpha g_nu g_beta g_lambda}
 *   G^II  = -1/4 delta_{beta lambda} Tr{Gam g_mu g_alpha g_nu}
 *   G^III = -1/4 delta_{alpha mu}    Tr{Gam g_nu g_beta g_lambda}
 * and, on the standard path, T^I = Vv, T^II_{abd} = Txv[b][a][d] + 1/4 d_bd Sxv[a],
 * T^III_{abd} = Txv[a][b][d] + Tyv[a][b][d] + 1/4 d_ad (Sxv + Syv)[b].
 * Traces are evaluated with the recursive delta-pairing rule (1/4 Tr of a product of even length
 * n = sum over partners of the last index, reference pi_pert.c:138-157 uses the same identity).
 *
 * Inputs with an index symmetry (Txv, Tyv symmetric in (a,d); T^II in (b,d); T^III in (a,d)) are
 * folded onto their a<=d (b<=d) representative, which shortens every row.
 */
#include <stdlib.h>
#include <string.h>

#include "kqed_assembly.h"

/* 1/4 Tr(g_i0 ... g_i{n-1}) by expanding the last gamma against each earlier one */
static int qtrace(int n, const int *idx);
int kqed_quarter_trace(int n, const int *idx) { return n >= 0 && n <= 16 ? qtrace(n, idx) : 0; }
static int qtrace(int n, const int *idx)
{
  if (n & 1) return 0;
  if (n == 0) return 1;
  int red[16], it = 0;
  for (int i = 0; i < n - 1; i++) {
    if (idx[i] != idx[n - 1]) continue;
    int k = 0;
    for (int j = 0; j < n - 1; j++) if (j != i) red[k++] = idx[j];
    it += ((i & 1) ? -1 : 1) * qtrace(n - 2, red);
  }
  return it;
}

/* Tr{Gam_{d r s} g_a ...}: Gam = 1/2 (g_d g_r g_s - g_d g_s g_r + g_r g_s g_d - g_s g_r g_d) */
static double gtrace(int d, int r, int s, int n, const int *rest)
{
  int seq[16];
  const int lead[4][3] = { { d, r, s }, { d, s, r }, { r, s, d }, { s, r, d } };
  const int sg[4] = { 1, -1, 1, -1 };
  double acc = 0.0;
  for (int t = 0; t < 4; t++) {
    memcpy(seq, lead[t], sizeof(int) * 3);
    memcpy(seq + 3, rest, sizeof(int) * n);
    acc += sg[t] * 4.0 * qtrace(3 + n, seq);
  }
  return 0.5 * acc;
}

static const int RS[6][2] = { { 0, 1 }, { 0, 2 }, { 0, 3 }, { 1, 2 }, { 1, 3 }, { 2, 3 } };

void kqed_dense_maps(double *Gfull /* [3][384][64] */)
{
  memset(Gfull, 0, sizeof(double) * 3 * 384 * 64);
  for (int rs = 0; rs < 6; rs++)
    for (int de = 0; de < 4; de++)
      for (int mu = 0; mu < 4; mu++)
        for (int nu = 0; nu < 4; nu++)
          for (int la = 0; la < 4; la++)
            for (int al = 0; al < 4; al++)
              for (int be = 0; be < 4; be++) {
                const int row = 64 * rs + 16 * mu + 4 * nu + la, col = 16 * al + 4 * be + de;
                const int r = RS[rs][0], s = RS[rs][1];
                const int i5[5] = { mu, al, nu, be, la };
                Gfull[0 * 384 * 64 + 64 * row + col] = gtrace(de, r, s, 5, i5) / 8.0;
                if (be == la) {
                  const int i3[3] = { mu, al, nu };
                  Gfull[1 * 384 * 64 + 64 * row + col] = -gtrace(de, r, s, 3, i3) / 4.0;
                }
                if (al == mu) {
                  const int i3[3] = { nu, be, la };
                  Gfull[2 * 384 * 64 + 64 * row + col] = -gtrace(de, r, s, 3, i3) / 4.0;
                }
              }
}

/* fold column c = 16a+4b+d onto the representative of its symmetry class */
static int fold(int c, int sym)
{
  int a = c >> 4, b = (c >> 2) & 3, d = c & 3;
  if (sym == 1 && a > d) { int t = a; a = d; d = t; }          /* (a,d) */
  if (sym == 2 && b > d) { int t = b; b = d; d = t; }          /* (b,d) */
  return 16 * a + 4 * b + d;
}

static void to_csr(const double *dense, int ncol, kqed_map *m)
{
  m->nnz = 0;
  m->kmax = 0;
  for (int r = 0; r < 384; r++) {
    m->rowptr[r] = m->nnz;
    for (int c = 0; c < ncol; c++) {
      if (dense[r * ncol + c] != 0.0) {
        m->col[m->nnz] = (short)c;
        m->val[m->nnz] = dense[r * ncol + c];
        m->nnz++;
      }
    }
    const int k = m->nnz - m->rowptr[r];
    if (k > m->kmax) m->kmax = k;
  }
  m->rowptr[384] = m->nnz;
}

int kqed_build_maps(kqed_map *std_map, kqed_map *eq0_map, kqed_map *eq0_mulam_map)
{
  double *G = (double *)malloc(sizeof(double) * 3 * 384 * 64);
  double *D = (double *)calloc(384 * 200, sizeof(double));
  if (!G || !D) { free(G); free(D); return 1; }
  kqed_dense_maps(G);
  const double *GI = G, *GII = G + 384 * 64, *GIII = G + 2 * 384 * 64;
  /* standard map over (Vv[64], Txv[64], Tyv[64], Sxv[4], Syv[4]) */
  for (int row = 0; row < 384; row++) {
    double *m = D + 200 * row;
    for (int c = 0; c < 64; c++) {
      const int al = c >> 4, be = (c >> 2) & 3, de = c & 3;
      m[c] += GI[64 * row + c];
      m[64 + fold(16 * be + 4 * al + de, 1)] += GII[64 * row + c];
      if (be == de) m[192 + al] += 0.25 * GII[64 * row + c];
      m[64 + fold(c, 1)] += GIII[64 * row + c];
      m[128 + fold(c, 1)] += GIII[64 * row + c];
      if (al == de) { m[192 + be] += 0.25 * GIII[64 * row + c]; m[196 + be] += 0.25 * GIII[64 * row + c]; }
    }
  }
  to_csr(D, 200, std_map);
  /* Tabd map over (T^I[64], T^II[64] sym (b,d), T^III[64] sym (a,d)) */
  memset(D, 0, sizeof(double) * 384 * 200);
  for (int row = 0; row < 384; row++) {
    double *m = D + 192 * row;
    for (int c = 0; c < 64; c++) {
      m[c] += GI[64 * row + c];
      m[64 + fold(c, 2)] += GII[64 * row + c];
      m[128 + fold(c, 1)] += GIII[64 * row + c];
    }
  }
  to_csr(D, 192, eq0_map);
  /* x = y: L(x,x) = -L_{mu<->lambda}(x,0) (reference kernels.c:30-48, 94-109) */
  double *E = (double *)malloc(sizeof(double) * 384 * 192);
  for (int i = 0; i < 384; i++) {
    const int rs = i / 64, mu = (i / 16) & 3, nu = (i / 4) & 3, la = i & 3;
    const int src = 64 * rs + 16 * la + 4 * nu + mu;
    for (int c = 0; c < 192; c++) E[192 * i + c] = -D[192 * src + c];
  }
  to_csr(E, 192, eq0_mulam_map);
  free(E); free(G); free(D);
  return 0;
}

void kqed_fill_g8(int *G8)
{
  /* 4 * (1/4 Tr) of eight gammas, index order (delta, rho, sigma, mu, alpha, nu, beta, lambda)
   * with i = sum idx_k 4^(7-k), as the reference's G8 table (pi_pert.c:161-177) */
  for (int i = 0; i < 65536; i++) {
    int idx[8];
    for (int k = 0; k < 8; k++) idx[k] = (i >> (2 * (7 - k))) & 3;
    G8[i] = 4 * qtrace(8, idx);
  }
}
