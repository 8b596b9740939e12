/*
 * kqed_oracle.h -- CPU oracle for the position-space QED kernel L_[rho,sigma];mu,nu,lambda(x,y).
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference library's FP64 CPU
 * evaluation (the AVX+FMA build of the src/LIBFILES sources), used as the checker for the HIP path.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the product
 * library (kqed_amd/) never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Floating-point evaluation follows
 * the reference's expression order; it is compiled with -ffp-contract=off so that every fused
 * multiply-add is one the reference writes explicitly (_mm256_fmadd_pd/_mm256_fmsub_pd).
 */
#ifndef KQED_ORACLE_H
#define KQED_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Table container: the reference's struct Grid_coeffs (src/HEADERS/KQED.h:69-87), flattened. */
typedef struct {
  int nstpx, nstpy, Nffa, nfmax;
  double *XX, *YY;
  double xstp, ystp;
  int *nfx;                 /* [nstpx] */
  float *Ffm, *Ffp;         /* [Nffa][nstpx][nstpy][nfmax], zero padded beyond nfx[ix] */
  double TX[23][810];       /* TAYLORX rows, KQED.h:57 */
  double TY[14][100];       /* TAYLORY rows, KQED.h:60 */
  int NY_tay, NX_tay;
  double *LUT[7];           /* ratio look-up tables, simd.c:716-784; index 1..6 */
  int nLUT;
} oq_tables;

/* CRC-32C and the DML rotate-xor accumulation (src/UTILS/crc32c.c:160-172). */
uint32_t oq_crc32c(uint32_t crc, const void *buf, size_t len);
void oq_checksum_accum(uint32_t *a, uint32_t *b, uint32_t rank, const void *buf, size_t len);

/* Load FFxy_single_cksum.bin, taylorx_cksum.bin, taylory_cksum.bin from dir (io.c:42-288).
 * Returns 0 on success, 1 on a missing file, bad magic, bad dimension or checksum mismatch. */
int oq_load(const char *dir, oq_tables *T);
void oq_free(oq_tables *T);

/* Public kernels (src/LIBFILES/kernels.c:82-231).  k has 384 doubles, flat index
 * 64*rhosig + 16*mu + 4*nu + lambda. */
void oq_L0(const oq_tables *T, const double x[4], const double y[4], double k[384]);
void oq_L1(const oq_tables *T, const double x[4], const double y[4], double k[384]);
void oq_L2(const oq_tables *T, const double x[4], const double y[4], double k[384]);
void oq_L3(const oq_tables *T, const double x[4], const double y[4], double k[384]);
void oq_ML2(const oq_tables *T, double M, const double x[4], const double y[4], double k[384]);

/* Batched driver (OpenMP over pairs) for the CPU baseline: which = 0..3 for L0..L3, 4 = M-L2. */
void oq_batch(const oq_tables *T, int which, double M, size_t n, const double *x, const double *y,
              double *out, int nthreads);
/* contraction with per-site FP32 tensors (pattern of amu_forlattice.c:99-107), s[ny], a[ny] = sum |terms| */
void oq_contract(const oq_tables *T, int which, double M, size_t nx, const double *x, const float *Tc, size_t ny,
                 const double *y, double *s, double *a, int nthreads);

/* Kernel combinations (all_kernels.c, con_kernel.c, sub_kernel.c, SYMXY.c, SYMXY0.c).  Output per
 * pair in the reference's struct layout: QED_Kernels (8 x 384: L0.xy, L0.yx, ..., L3.yx), Kernels
 * (4 x 384: L0..L3) or Kernels[3] (12 x 384).  M has 4 entries (NULL = zeros); returns 1 on a bad
 * comb. */
enum {
  OQ_COMB_ALL = 0,      /* compute_all_kernels            -> QED_Kernels */
  OQ_COMB_ALL_M = 1,    /* compute_all_Mkernels           -> QED_Kernels */
  OQ_COMB_ALL_M_V2 = 2, /* compute_all_Mkernels_v2        -> QED_Kernels */
  OQ_COMB_CON = 3,      /* compute_con_kernels            -> QED_Kernels */
  OQ_COMB_CON_V2 = 4,   /* compute_con_kernels_v2         -> Kernels[3] */
  OQ_COMB_CON_M_L2 = 5, /* compute_con_kernelsM_L2        -> Kernels[3] */
  OQ_COMB_SUB_M_L2 = 6, /* compute_sub_kernelsM_L2        -> Kernels[3] */
  OQ_COMB_SYMXY = 7,    /* compute_all_kernels_SYMXY      -> QED_Kernels */
  OQ_COMB_SYMXY_V2 = 8, /* compute_all_kernels_SYMXY_v2   -> Kernels */
  OQ_COMB_SYMXY0 = 9,   /* compute_all_kernels_SYMXY0     -> QED_Kernels */
  OQ_COMB_SYMXY0_V2 = 10 /* compute_all_kernels_SYMXY0_v2 -> Kernels */
};
int oq_comb_len(int comb); /* doubles per pair */
int oq_combine(const oq_tables *T, int comb, const double M[4], const double x[4], const double y[4], double *K);
void oq_combine_batch(const oq_tables *T, int comb, const double M[4], size_t n, const double *x, const double *y,
                      double *out, int nthreads);
/* s[j*nk + q] = sum_i K_q(x_i, y_j) . Tc[i] for the nk output kernels of comb; a: sum of |terms| */
void oq_contract_comb(const oq_tables *T, int comb, const double M[4], size_t nx, const double *x, const float *Tc,
                      size_t ny, const double *y, double *s, double *a, int nthreads);

/* Perturbative lepton loop (pi_pert.c).  x and y are in units of the loop lepton's inverse mass
 * (the reference's example1 passes M_v x, M_v y).  Singular where x, y or x-y vanishes. */
int oq_diractrace(int n, const int *idx);                     /* pi_pert.c:136-157 */
void oq_bessk(int n, double x, double *bv);                   /* pi_pert.c:92-133: K_0..K_n */
void oq_pihat1(const double x[4], const double y[4], double *v1 /* [4][4][4][4][4] */,
               double *r1 /* [4][4][4][4] */);                  /* pi_pert.c:222-381 */
void oq_ipihat(const double x[4], const double y[4], double *out /* [6][4][4][4] */); /* :180-220 */
void oq_ipihat_batch(size_t n, const double *x, const double *y, double *out, int nthreads);

/* Pieces exposed for unit tests. */
double oq_chebU(int order, int nk, double x, const double *co);          /* cheby.c:13-91 */
double oq_chebU_avx(int order, int nk, double x, const double *co, const oq_tables *T); /* simd.c:182-336 */
int oq_bsrch(const double *arr, double target, int lo, int hi);         /* getff-new.c:199-210 */
/* S/T/V tensors of the standard path (init_STV, QED_kernel.c:427-520); returns 1 when out of grid.
 * stv layout: Vv[64], Txv[64], Tyv[64], Sxv[4], Syv[4]. */
int oq_stv(const oq_tables *T, const double x[4], const double y[4], double stv[200]);
/* Tabd_xeq0 / Tabd_yeq0 (Tabd.c:23-316): t = tI[64], tII[64], tIII[64]. */
int oq_tabd_xeq0(const oq_tables *T, const double y[4], double t[192]);
int oq_tabd_yeq0(const oq_tables *T, const double x[4], double t[192]);
/* Dense assembly maps (384 x 200 and 384 x 192) built from the Dirac traces; row-major. */
void oq_assembly_maps(double *Mstd, double *Meq0);
/* The 32(+2) interpolated weight-function scalars of the standard regime, in this order:
 * QG0{dcb,dx,dy,v}, QL2{4}, QL4{4}, QG1{4}, QG3dy, QG2(d0){4}, dxQG2{4}, dxQG3{4}, QG2(d1){4},
 * QG3(d1){4}, swap: fq1,fq2(QG2), fq1,fq2(QG3).  Returns the regime code (see oq_regime). */
int oq_interp_scalars(const oq_tables *T, const double x[4], const double y[4], double s[45]);
/* Regime of a pair: 0 zero, 1 x=0, 2 y=0, 3 x=y, 4 Taylor-x, 5 standard, 6 out of grid;
 * +8 when the x<->x-y swap is active. */
int oq_regime(const oq_tables *T, const double x[4], const double y[4]);

/* S/T/V tensors (layout of oq_stv) and L0 of a standard-regime pair (XX[0] <= |x| <= XX[last])
 * with the interpolated weight functions taken from w instead of the tables: the chain-rule
 * algebra of chnr_dS.c:50-95, chnr_dT.c:185-293, chnr_dV.c:363-479 (swap included) on given
 * inputs.  w, 34 values in the order of the device library's kqed_batch_inject:
 *   g0: dcb dx dy | l2: dx dy val dcb | l4: dx dy val dcb | v1: dx dy val dcb |
 *   g2: val dy | dcb dydcb dcbdcb | dxdx dx dxdy dxdcb | g3: dxdx dx dxdy dxdcb |
 *   dcb dydcb dcbdcb | dy | swap only: g2 dydy, g3 dydy
 * (the last two replace the finite differences of chnr_dV.c:444-466).  Returns 1 as oq_stv. */
int oq_stv_inject(const oq_tables *T, const double x[4], const double y[4], const double w[34], double stv[200]);
void oq_L0_inject(const oq_tables *T, const double x[4], const double y[4], const double w[34], double k[384]);

/* The g1/g2/g3 derivatives chnr_dV hands to its tensor step (struct vtmp of chnr_dV.c, in field
 * order: g2, dg1dx, dg1dy, dg1dcb, dg2dx, dg2dy, dg2dcb, dg3dx, dg3dy, dg3dcb, ddg2dxdx, ddg2dxdy,
 * ddg2dxdcb, ddg2dydcb, ddg2dcbdcb, ddg3dxdx, ddg3dxdy, ddg3dxdcb, ddg3dydcb, ddg3dcbdcb,
 * ddg1dxdx, ddg1dxdy, ddg1dxdcb, ddg1dydcb, ddg1dcbdcb, ddg1dydy, ddg2dydy), after the swap. */
int oq_vterms(const oq_tables *T, const double x[4], const double y[4], double v[27]);

/* Deliberate mistakes for tests that show a pin catches them (0 switches them off): 1 chnr_dS c2
 * sign, 2 chnr_dT dell3dx term dropped, 3 chnr_dT t2 index, 4 swap-T chain-rule sign, 5 chnr_dV
 * ddg1dxdy term dropped, 6 get_dv operand transposed, 7 swap-V chain-rule sign, 8 swap-S factor,
 * 9 Taylor-x dell3dx term dropped, 10 Taylor-x dg0dx blend reversed, 11 x-Hermite weights
 * exchanged, 12 d/dy array exponent, 13 Clenshaw ratio table offset, 14 Taylor-x ddg1dxdy term
 * dropped. */
void oq_plant(int k);

/* Algorithmic FP64 flop count of the last oq_L0 calls on this thread (reference algorithm). */
double oq_flops_get(void);
double oq_flops_build_get(void); /* part spent on per-pair coefficient arrays (node-only work) */
void oq_flops_reset(void);

#ifdef __cplusplus
}
#endif
#endif
