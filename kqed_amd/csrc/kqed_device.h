/*
 * kqed_device.h -- device-side data layout shared by the HIP kernels and the host API.
 *
 * HBM layout of the tables (all built once by initialise()):
 *   P   [14][nstpx][nstpy][ks] double2 : Clenshaw inputs of the packed (vector) interpolation path,
 *                                        .x = value array, .y = d/dy array, per grid node
 *   PNC [4][nstpx][nstpy][ks]  double  : d/dy arrays of the derivative-only variant for the tables
 *                                        QG2, dxQG2, QG3, dxQG3
 *   S   [14][nstpx][nstpy][ks] double2 : Clenshaw inputs of the scalar interpolation path (Taylor
 *                                        region and grid edges), .x = value, .y = d/dy
 *   LUT [ks][8] double                 : Clenshaw derivative ratios, slot 2d-1/2d for order d
 * ks = nfmax + 2 + KQED_CLEN_BLOCK - 1: the longest Clenshaw (nfx + mm, mm <= 2) rounded up to whole
 * blocks of KQED_CLEN_BLOCK steps; the slots past a node's length are zero.
 *
 * Per-pair record written by the interpolation kernel and read by the assembly kernel:
 *   rec[REC_LEN] doubles: x[4], y[4], coef[4], M, gaussX, gaussY, input index, main payload [64],
 *   three Tabd payloads [20], pad
 *   kinds (int): 3 bits per evaluation slot (KIND_*)
 */
#ifndef KQED_DEVICE_H
#define KQED_DEVICE_H

#include <hip/hip_runtime.h>

#define KQED_REC_LEN 144
#define KQED_REC_X 0      /* x[4] */
#define KQED_REC_Y 4      /* y[4] */
#define KQED_REC_COEF 8   /* coefficient of each evaluation slot [4] */
#define KQED_REC_M 12     /* M of M-L2 */
#define KQED_REC_GX 13    /* exp(-M x^2 / 2) */
#define KQED_REC_GY 14    /* exp(-M y^2 / 2) */
#define KQED_CLEN_BLOCK 8  /* Clenshaw steps per block of coefficient loads (at most) */
#define KQED_REC_P 15     /* input index of the pair */
#define KQED_REC_MAIN 16  /* slot 0 payload [64]: stage-1 record, then the 36 primitives */
#define KQED_REC_AUX 80   /* slots 1..3 payloads [20] each */
#define KQED_HELP_BASE 36 /* main payload: after the 36 primitives, the assembly's helper inputs [26] */
#define KQED_AUX_LEN 20

/* Learned surrogate of the packed-path interpolation: the interpolated weight functions it stands
 * in for, as double offsets into the stage-1 record (checked against the struct in the kernels):
 * dS 3, dT 12, dV 17, and the two swap-regime second y derivatives. */
#define KQED_NSURR 34
#define KQED_SURR_SLOTS                                                                     \
  { 14, 12, 13, 23, 24, 16, 25, 33, 32, 19, 34, 20, 21, 18, 22, 35, 40,                     \
    41, 48, 49, 45, 39, 46, 47, 50, 42, 51, 52, 44, 53, 54, 43, 61, 60 }
#define KQED_S1_FLAG2 10    /* stage-1 record: swap flag */
#define KQED_S1_NEEDALG 11  /* stage-1 record: chain-rule algebra still to run */
#define KQED_SURR_BIT 0x80000000u /* kinds bit: slot 0 interpolated by the surrogate */
#define KQED_SURR_NB 24           /* (log x, log y) bins of the output scales and of the error check */
#define KQED_SURR_YMIN 1e-3       /* lower end of the y range of the bins */
#define KQED_SURR_ROW 40          /* doubles per row of extracted training data (5 + 34, padded) */

/* Entries per lane of the six row slots of every wave-dealt assembly map (rows sorted by length;
 * the standard map needs exactly these, the Tabd maps fewer and are padded with the zero slot) */
#define KQED_WM_K { 13, 9, 6, 6, 4, 3 }
#define KQED_WM_PER_LANE 41
#define KQED_WM_MAX (3 * 64 * KQED_WM_PER_LANE) /* u32 entries of the three maps */
#define KQED_LDS_ARRAYS 8       /* node arrays per wavefront staged in LDS by the packed-path kernel */
#define KQED_INTERP_THREADS 128 /* interpolation block: two wavefronts (256 measured 1.5-2% slower) */
#define KQED_ASM2_WAVES 12  /* wavefronts per two-pair assembly block (LDS 100 KB; 16 measured 18 % slower: VGPR-bound) */
#ifndef KQED_ASM2D_WAVES
#define KQED_ASM2D_WAVES 12 /* the contraction shortcut's dot form of the two-pair kernel (no map in LDS; 16 measured 14 % slower: spills) */
#endif
#define KQED_ASM2M_WAVES 8 /* multi-slot two-pair assembly: LDS 139 KB (all three maps, record, O) */
#define KQED_ASM2T_WAVES 12 /* L1-L3 two-pair assembly with one Tabd map application: LDS 112 KB */
#define KQED_ASM_WAVES 16   /* wavefronts per assembly block: one block (LDS 136 KB) fills a CU */

/* Pulled-back site weights of the contraction shortcut hold only the standard map's columns that
 * some row reads: the map reads the symmetric T entries (Txv, Tyv: columns 64..191, index alpha beta
 * delta) only at alpha <= delta, so 152 of the 200 columns.  kqed_wcol: compact index of column c,
 * or -1 for a column no row reads. */
#define KQED_WCOLS 152
/* the pull-back deals these columns, longest first, to three slots of 64 lanes, each unrolled to
 * its longest column: 26, 18 and 4 entries (checked against the map at init) */
#define KQED_PB_K { 26, 18, 4 }
#define KQED_PB_TOTAL 48
__host__ __device__ inline int kqed_wcol(int c)
{
  if (c < 64) return c;
  if (c >= 192) return c - 48;
  const int t = (c - 64) >> 6, a = (c >> 4) & 3, b = (c >> 2) & 3, d = c & 3;
  if (a > d) return -1;
  /* (a, d) with a <= d in the order (0,0) (0,1) (0,2) (0,3) (1,1) (1,2) (1,3) (2,2) (2,3) (3,3) */
  const int ad = a * 4 - a * (a - 1) / 2 + (d - a);
  return 64 + 40 * t + 10 * b + ad;
}

enum {
  KIND_NONE = 0,
  KIND_STD = 1,       /* standard / Taylor / swap path: S, T, V tensors (200 inputs) */
  KIND_XEQ0 = 2,      /* L0(0, v): Tabd_xeq0 tensors (192 inputs) */
  KIND_YEQ0 = 3,      /* L0(v, 0): Tabd_yeq0 tensors */
  KIND_YEQ0_MULAM = 4,/* L0(v, v) = -L0_{mu<->lambda}(v, 0) */
  KIND_XEQ0_M = 5,    /* M-L2 subtraction term built on L0(0, y) */
  KIND_YEQ0_M = 6     /* M-L2 subtraction term built on L0(x, 0) */
};

struct kqed_dev_tables {
  int nstpx, nstpy, nfmax, ks;
  const double *XX, *YY;
  const int *nfx;
  const double *TX;     /* [23][810] */
  const double *TY;     /* [14][100] */
  const double *LUT;    /* [ks][8] */
  const double2 *P;
  const double *PNC;
  const double2 *S;
  /* assembly maps (0 standard, 1 Tabd, 2 Tabd with mu<->lambda and sign) dealt to the 64 lanes of
   * a wavefront: rows sorted by length, row 64m + l of the sorted order goes to lane l, slot m;
   * entries [m][k][lane] as u32 = col * 16 | (upper 16 bits of the FP64 coefficient) << 16 with
   * coefficient +-0.5, +-1 or +-2; padded entries read the zero slot (col 200) */
  const unsigned *wmE;         /* all three maps, concatenated */
  const unsigned short *wmRow; /* [3][6][64] output row of (map, slot, lane) */
  /* LDS slot of each of the standard map's 200 tensor entries (its column c is read from slot
   * wmSlot[c]; the zero column stays at 200); the Tabd maps use their columns as slots */
  const unsigned char *wmSlot;
  int wmOff[3][7];             /* start of slot m of map q in wmE; wmOff[q][6] = end */
  int wmK[3][6];               /* entries per lane in slot m of map q */
  int wmTotal;
};

#endif
