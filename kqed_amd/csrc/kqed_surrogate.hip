/*
 * kqed_surrogate.hip -- the learned surrogate of the packed-path interpolation: an MLP
 * (3 -> 128 -> 128 -> 128 -> 128 -> 34, SiLU) with BF16 operands and FP32 accumulation on the
 * matrix cores (v_mfma_f32_16x16x16_bf16), its training kernels (GEMM, SiLU forward/backward, loss,
 * Adam) and the fused inference kernel that writes the 34 weight functions into stage-1 records.
 *
 * The network stands in for the table interpolation of reference getff-new.c:214-312 /
 * simd.c:343-687 (the weight functions and derivatives consumed by chnr_dS.c, chnr_dT.c,
 * chnr_dV.c); the chain-rule algebra and the assembly that follow are unchanged.
 *
 * MFMA 16x16x16 operand layout (lane l, element j = 0..3):
 *   A (M x K): A[l % 16][4 (l / 16) + j]    B (K x N): B[4 (l / 16) + j][l % 16]
 *   D (M x N): D[4 (l / 16) + j][l % 16]
 */
#include <hip/hip_runtime.h>

#include "kqed_device.h"

#define DEV __device__ __forceinline__

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

/* FP32 -> BF16, round to nearest even (finite inputs) */
DEV unsigned short f2bf(float f)
{
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
DEV float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

/* ---------------------------------------------------------------- training GEMM */
/* C[z][m][n] = sum_{k in split z} A(m,k) B(k,n) (+ bias[n] on split 0), A(m,k) = A[m*sam + k*sak],
 * B(k,n) = B[k*sbk + n*sbn]; operands rounded to BF16, FP32 accumulation.  One wavefront per 32x32
 * tile of C (2x2 MFMA tiles); grid (ceil(N/32), ceil(M/32), splits), split z covers
 * k in [z*kper, min(K, (z+1)*kper)), kper a multiple of 16. */
__global__ void __launch_bounds__(64)
k_gemm_bf16(int M, int N, int K, int kper, const float *__restrict__ A, long sam, long sak, const float *__restrict__ B,
            long sbk, long sbn, float *__restrict__ C, long ldc, long zstride, const float *__restrict__ bias)
{
  const int lane = threadIdx.x, r = lane & 15, q = lane >> 4;
  const int m0 = blockIdx.y * 32, n0 = blockIdx.x * 32, z = blockIdx.z;
  const int k0 = z * kper, k1 = min(K, k0 + kper);
  f32x4 acc[2][2];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) acc[a][b] = f32x4{ 0.f, 0.f, 0.f, 0.f };
  for (int kk = k0; kk < k1; kk += 16) {
    bf16x4 fa[2], fb[2];
    for (int t = 0; t < 2; t++) {
      const int i = m0 + 16 * t + r, n = n0 + 16 * t + r;
      for (int j = 0; j < 4; j++) {
        const int k = kk + 4 * q + j;
        const float va = (i < M && k < k1) ? A[(long)i * sam + (long)k * sak] : 0.f;
        const float vb = (n < N && k < k1) ? B[(long)k * sbk + (long)n * sbn] : 0.f;
        fa[t][j] = (short)f2bf(va);
        fb[t][j] = (short)f2bf(vb);
      }
    }
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fa[a], fb[b], acc[a][b], 0, 0, 0);
  }
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int j = 0; j < 4; j++) {
        const int i = m0 + 16 * a + 4 * q + j, n = n0 + 16 * b + r;
        if (i < M && n < N) C[(long)z * zstride + (long)i * ldc + n] = acc[a][b][j] + ((bias && z == 0) ? bias[n] : 0.f);
      }
}

/* out[i] = sum_z part[z][i] (split-K partial sums, fixed order) */
__global__ void k_sum_splits(long n, int nz, long zstride, const float *__restrict__ part, float *__restrict__ out)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int z = 0; z < nz; z++) s += part[(long)z * zstride + i];
  out[i] = s;
}

DEV float sigm(float z) { return 1.f / (1.f + __expf(-z)); }

__global__ void k_silu(long n, const float *__restrict__ Z, float *__restrict__ A)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) A[i] = Z[i] * sigm(Z[i]);
}

/* dZ = dA * silu'(Z) */
__global__ void k_silu_bwd(long n, const float *__restrict__ Z, const float *__restrict__ dA, float *__restrict__ dZ)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float s = sigm(Z[i]);
  dZ[i] = dA[i] * s * (1.f + Z[i] * (1.f - s));
}

/* db[n] = sum_b dZ[b][n] in a fixed order: a block of 1024 threads takes 64 columns; thread
 * (phase p = t / 64, column t % 64) sums rows b = p, p + 16, ... (a wavefront reads 64 consecutive
 * columns of one row), then the 16 phases are added in order */
__global__ void __launch_bounds__(1024) k_colsum_f(int B, int N, const float *__restrict__ dZ, float *__restrict__ db)
{
  __shared__ float ps[16][64];
  const int c = threadIdx.x & 63, p = threadIdx.x >> 6;
  const int n = blockIdx.x * 64 + c;
  float s = 0.f;
  if (n < N)
    for (int b = p; b < B; b += 16) s += dZ[(long)b * N + n];
  ps[p][c] = s;
  __syncthreads();
  if (p == 0 && n < N) {
    float t = 0.f;
    for (int k = 0; k < 16; k++) t += ps[k][c];
    db[n] = t;
  }
}

/* Mean squared error over the valid (sample, output) entries: pred[B][NP], tgt/wgt[B][NOUT].
 * dpred = 2 w (pred - tgt) / norm; per-block partial sums of w (pred - tgt)^2 in part[block]. */
__global__ void k_mse(int B, int NP, int NOUT, const float *__restrict__ pred, const float *__restrict__ tgt,
                      const float *__restrict__ wgt, float norm, float *__restrict__ dpred, float *__restrict__ part)
{
  __shared__ float red[256];
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (i < (long)B * NP) {
    const long b = i / NP;
    const int o = (int)(i % NP);
    float g = 0.f;
    if (o < NOUT) {
      const float w = wgt[b * NOUT + o], d = pred[i] - tgt[b * NOUT + o];
      l = w * d * d;
      g = 2.f * w * d / norm;
    }
    dpred[i] = g;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

/* Adam (bias-corrected step size lr_t precomputed on the host) */
__global__ void k_adam(long n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr_t, float b1, float b2, float eps)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  m[i] = b1 * m[i] + (1.f - b1) * gi;
  v[i] = b2 * v[i] + (1.f - b2) * gi * gi;
  p[i] -= lr_t * m[i] / (sqrtf(v[i]) + eps);
}

/* ---------------------------------------------------------------- normalisation */
/* network inputs from the post-swap invariants: u (log x), v (log y), cb, each in [-1, 1] */
struct SurrRange { float lx0, lx1, ly0, ly1; };

DEV void surr_inputs(const SurrRange &R, double x, double y, double cb, float in[3], float &bu, float &bv)
{
  const float u = ((float)log(x) - R.lx0) / (R.lx1 - R.lx0), v = ((float)log(y) - R.ly0) / (R.ly1 - R.ly0);
  in[0] = 2.f * u - 1.f;
  in[1] = 2.f * v - 1.f;
  in[2] = (float)cb;
  bu = fminf(fmaxf(u * KQED_SURR_NB - 0.5f, 0.f), KQED_SURR_NB - 1.f);
  bv = fminf(fmaxf(v * KQED_SURR_NB - 0.5f, 0.f), KQED_SURR_NB - 1.f);
}

/* output scale of weight function k at continuous bin coordinates (bu, bv): exp of the bilinear
 * interpolation of the per-bin log maxima lsc[k][NB][NB] */
DEV float surr_scale(const float *__restrict__ lsc, int k, float bu, float bv)
{
  const int iu = min((int)bu, KQED_SURR_NB - 2), iv = min((int)bv, KQED_SURR_NB - 2);
  const float fu = bu - iu, fv = bv - iv;
  const float *g = lsc + (long)k * KQED_SURR_NB * KQED_SURR_NB;
  const float a = g[iu * KQED_SURR_NB + iv], b = g[iu * KQED_SURR_NB + iv + 1];
  const float c = g[(iu + 1) * KQED_SURR_NB + iv], d = g[(iu + 1) * KQED_SURR_NB + iv + 1];
  return __expf((1.f - fu) * ((1.f - fv) * a + fv * b) + fu * ((1.f - fv) * c + fv * d));
}

/* Training batch: rows idx[b] of the extracted data (KQED_SURR_ROW doubles each) -> inputs
 * X[b][16] (3 used), normalised targets T[b][34] and weights W[b][34] (0 for rows that are not
 * packed-path pairs, and for the two swap outputs of non-swap rows). */
__global__ void k_surr_batch(int B, const unsigned *__restrict__ idx, const double *__restrict__ data, SurrRange R,
                             const float *__restrict__ lsc, float *__restrict__ X, float *__restrict__ Tg,
                             float *__restrict__ W)
{
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const double *row = data + (long)idx[b] * KQED_SURR_ROW;
  const bool ok = row[4] != 0.0;
  float in[3], bu = 0.f, bv = 0.f;
  if (ok) surr_inputs(R, row[0], row[1], row[2], in, bu, bv);
  else in[0] = in[1] = in[2] = 0.f;
  for (int k = 0; k < 16; k++) X[(long)b * 16 + k] = k < 3 ? in[k] : 0.f;
  const bool swap = row[3] != 0.0;
  for (int k = 0; k < KQED_NSURR; k++) {
    const bool use = ok && (k < 32 || swap);
    Tg[(long)b * KQED_NSURR + k] = use ? (float)(row[5 + k] / (double)surr_scale(lsc, k, bu, bv)) : 0.f;
    W[(long)b * KQED_NSURR + k] = use ? 1.f : 0.f;
  }
}

/* per-bin maxima of |f_k| over the packed rows (as float bits: non-negative floats order as ints) */
__global__ void k_surr_binmax(long n, const double *__restrict__ data, SurrRange R, unsigned *__restrict__ bmax,
                              unsigned *__restrict__ bcount)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *row = data + i * KQED_SURR_ROW;
  if (row[4] == 0.0) return;
  float in[3], bu, bv;
  surr_inputs(R, row[0], row[1], row[2], in, bu, bv);
  const int cell = (int)(bu + 0.5f) * KQED_SURR_NB + (int)(bv + 0.5f);
  atomicAdd(&bcount[cell], 1u);
  for (int k = 0; k < KQED_NSURR; k++) {
    const float a = fabsf((float)row[5 + k]);
    if (a > 0.f && isfinite(a)) atomicMax(&bmax[k * KQED_SURR_NB * KQED_SURR_NB + cell], __float_as_uint(a));
  }
}

/* ---------------------------------------------------------------- inference */
/* Packed inference weights: layer l as BF16 in fragment order Wp[k/4][n][k%4] (the four K values
 * of one MFMA B fragment are contiguous), K_l = {16, 128, 128, 128, 128}, N_l = {128, 128, 128,
 * 128, 48}; FP32 biases. */
struct SurrNet {
  const unsigned short *W[5];
  const float *b[5];
  const float *lsc;
  SurrRange R;
};

#define SW 4 /* wavefronts per inference block, 16 records each */

/* One layer for the 16 rows of a wavefront: out = act(in[16][K] W + b), in/out BF16 in LDS
 * ([16][128] row-major), or the FP32 result of the last layer into fo[16][48]. */
template <int K, int N, bool LAST>
DEV void surr_layer(const unsigned short *__restrict__ Wp, const float *__restrict__ bias, const unsigned short *in,
                    unsigned short *out, float *fo, int lane)
{
  const int r = lane & 15, q = lane >> 4;
  f32x4 acc[N / 16];
#pragma unroll
  for (int t = 0; t < N / 16; t++) acc[t] = f32x4{ 0.f, 0.f, 0.f, 0.f };
#pragma unroll
  for (int kk = 0; kk < K; kk += 16) {
    const bf16x4 a = *(const bf16x4 *)(in + r * 128 + kk + 4 * q);
#pragma unroll
    for (int t = 0; t < N / 16; t++) {
      const bf16x4 b = *(const bf16x4 *)(Wp + (((long)(kk / 4 + q) * N + 16 * t + r) << 2));
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < N / 16; t++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int i = 4 * q + j, n = 16 * t + r;
      const float z = acc[t][j] + bias[n];
      if (LAST) fo[i * 48 + n] = z;
      else out[i * 128 + n] = f2bf(z * sigm(z));
    }
}

/* Records of one pass flagged KQED_SURR_BIT get their 34 weight functions from the network
 * (written as FP64 into the stage-1 slots).  Each wavefront runs 16 consecutive records. */
__global__ void __launch_bounds__(64 * SW)
k_surr_records(size_t n, double *__restrict__ rec, const unsigned *__restrict__ kinds, SurrNet net)
{
  __shared__ unsigned short sA[SW][2][16 * 128];
  __shared__ float sO[SW][16 * 48];
  __shared__ float sS[SW][16][2];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const size_t base = ((size_t)blockIdx.x * SW + w) * 16;
  if (base >= n) return;
  unsigned any = 0;
  if (lane < 16) {
    const size_t p = base + lane;
    float in[3] = { 0.f, 0.f, 0.f }, bu = 0.f, bv = 0.f;
    if (p < n && (kinds[p] & KQED_SURR_BIT)) {
      const double *m = rec + p * KQED_REC_LEN + KQED_REC_MAIN;
      surr_inputs(net.R, m[0], m[2], m[1], in, bu, bv); /* stage-1 order: x, cb, y */
      any = 1;
    }
    for (int k = 0; k < 16; k++) sA[w][0][lane * 128 + k] = f2bf(k < 3 ? in[k] : 0.f);
    sS[w][lane][0] = bu;
    sS[w][lane][1] = bv;
  }
  if (!__ballot(any)) return; /* wave-uniform */
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  surr_layer<16, 128, false>(net.W[0], net.b[0], sA[w][0], sA[w][1], nullptr, lane);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  surr_layer<128, 128, false>(net.W[1], net.b[1], sA[w][1], sA[w][0], nullptr, lane);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  surr_layer<128, 128, false>(net.W[2], net.b[2], sA[w][0], sA[w][1], nullptr, lane);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  surr_layer<128, 128, false>(net.W[3], net.b[3], sA[w][1], sA[w][0], nullptr, lane);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  surr_layer<128, 48, true>(net.W[4], net.b[4], sA[w][0], nullptr, sO[w], lane);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int slots[KQED_NSURR] = KQED_SURR_SLOTS;
  for (int e = lane; e < 16 * KQED_NSURR; e += 64) {
    const int i = e / KQED_NSURR, k = e % KQED_NSURR;
    const size_t p = base + i;
    if (p >= n || !(kinds[p] & KQED_SURR_BIT)) continue;
    double *m = rec + p * KQED_REC_LEN + KQED_REC_MAIN;
    const bool swap = m[KQED_S1_FLAG2] != 0.0;
    const float s = surr_scale(net.lsc, k, sS[w][i][0], sS[w][i][1]);
    m[slots[k]] = (k < 32 || swap) ? (double)(sO[w][i * 48 + k] * s) : 0.0;
  }
}

/* per-pair normwise error |a - b|_max / |b|_max of two FP64 kernel batches, accumulated per bin
 * of the pair (bins given per pair, -1 = not counted): max as float bits, count */
__global__ void k_surr_binerr(size_t n, const double *__restrict__ a, const double *__restrict__ b,
                              const int *__restrict__ bin, unsigned *__restrict__ emax, unsigned *__restrict__ ecount,
                              float *__restrict__ err)
{
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double num = 0.0, den = 0.0;
  for (int i = 0; i < 384; i++) {
    num = fmax(num, fabs(a[p * 384 + i] - b[p * 384 + i]));
    den = fmax(den, fabs(b[p * 384 + i]));
  }
  const float e = den > 0.0 ? (float)(num / den) : (num > 0.0 ? 1e30f : 0.f);
  err[p] = e;
  const int c = bin[p];
  if (c < 0) return;
  atomicAdd(&ecount[c], 1u);
  atomicMax(&emax[c], __float_as_uint(e));
}
