// Learned Davidson preconditioner: a per-determinant MLP (no counterpart in
// the reference, whose preconditioner is R/(E0-d), davidson.h:861-868).
//
//   x_I (160 features) -> Linear(160,128) -> GELU -> Linear(128,128) -> GELU
//   (GELU in tanh form, z * sigmoid(2u), u = sqrt(2/pi)(z + 0.044715 z^3))
//                      -> Linear(128,1) = c_I
// Weights are kept in FP32 (master copy, Adam state) and used in BF16; all
// products run on MFMA (training: v_mfma_f32_32x32x8_bf16; the solver's
// inference kernel: gfx950's v_mfma_f32_32x32x16_bf16) with FP32 accumulation.
// Activations entering a product are rounded to BF16 at fixed points (X, H1,
// H2, dZ1, dZ2, and Z1/Z2 where GELU' is evaluated), which the CPU gradient
// check in the tests reproduces.
//
// Features of determinant I = (Ia, Ib) at Ritz value E with residual r
// (unit norm over N determinants):
//   [0,64)   alpha occupation bits     [64,128) beta occupation bits
//   128      r_I*sqrt(N)               129      (h_II - E)/8
//   130      clamp(r_I*sqrt(N)/(E - h_II), +-50)/10   (the diagonal step)
//   131,132  alpha / beta excitation level vs string 0, /4
//   133      1                          [134,160) 0
// Target: the matching component of the unit error vector, times sqrt(N).
//
// MFMA operand layout used below (32x32x8, wave64):
//   A (32x8):  lane l holds A[l%32][4*(l/32) + j], j = 0..3
//   B (8x32):  lane l holds B[4*(l/32) + j][l%32]
//   D (32x32): lane l, register r holds D[8*(r/4) + 4*(l/32) + r%4][l%32]
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>

#include "common.h"

namespace sbdx {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kDin = 160, kHid = 128;
constexpr int kW1 = 0, kB1 = kW1 + kDin * kHid, kW2 = kB1 + kHid, kB2 = kW2 + kHid * kHid,
              kW3 = kB2 + kHid, kB3 = kW3 + kHid, kNParam = kB3 + 4;

// FP32 -> BF16, round to nearest even (one v_cvt_pk_bf16_f32 on gfx950)
__device__ __forceinline__ unsigned short f2bf(float f) { return __builtin_bit_cast(unsigned short, (__bf16)f); }
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float(((unsigned)h) << 16); }
// GELU in tanh form, 0.5 z (1 + tanh(u)) = z * sigmoid(2u) with
// u = sqrt(2/pi) (z + 0.044715 z^3): one exp and one reciprocal
constexpr float kGeluC = 0.79788456080286536f;  // sqrt(2/pi)
__device__ __forceinline__ float gelu_sig(float z) {
  const float u = kGeluC * z * (1.0f + 0.044715f * z * z);
  return __builtin_amdgcn_rcpf(1.0f + __expf(-2.0f * u));
}
__device__ __forceinline__ float gelu(float z) { return z * gelu_sig(z); }
__device__ __forceinline__ float gelu_d(float z) {
  const float s = gelu_sig(z);
  return s + z * s * (1.0f - s) * 2.0f * kGeluC * (1.0f + 3.0f * 0.044715f * z * z);
}
// The same GELU for the inference kernel, its constants folded:
// sigmoid(2u) = 1 / (1 + 2^t), t = z (k1 + k2 z^2), k1 = -2 sqrt(2/pi) log2(e),
// k2 = 0.044715 k1; native exp2 and reciprocal (v_exp_f32, v_rcp_f32).
constexpr float kGeluT1 = -2.0f * 0.79788456080286536f * 1.4426950408889634f;
constexpr float kGeluT2 = 0.044715f * kGeluT1;
__device__ __forceinline__ float gelu_fast(float z) {
  const float t = z * __builtin_fmaf(kGeluT2, z * z, kGeluT1);
  return z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(t));
}
__device__ __forceinline__ int d_row(int r, int lane) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

// Features of one determinant as BF16 into dst[0..kDin).
__device__ void features(unsigned short *dst, uint64_t a, uint64_t b, uint64_t a0, uint64_t b0,
                         double r_scaled, double hmE) {
  for (int k = 0; k < 64; k++) {
    dst[k] = f2bf((float)((a >> k) & 1ull));
    dst[64 + k] = f2bf((float)((b >> k) & 1ull));
  }
  double den = -hmE;  // E - h_II
  if (fabs(den) < 1e-3) den = den < 0 ? -1e-3 : 1e-3;
  double dstep = r_scaled / den;
  dstep = fmax(-50.0, fmin(50.0, dstep));
  dst[128] = f2bf((float)r_scaled);
  dst[129] = f2bf((float)(hmE / 8.0));
  dst[130] = f2bf((float)(dstep / 10.0));
  dst[131] = f2bf((float)(__popcll(a ^ a0) / 8.0));
  dst[132] = f2bf((float)(__popcll(b ^ b0) / 8.0));
  dst[133] = f2bf(1.0f);
  for (int k = 134; k < kDin; k++) dst[k] = 0;
}

// One 32-row tile: acc[nt] = A(32 x K) * B(K x 128) for the 4 column tiles.
// A row-major in LDS with leading dim lda; B row-major [K][128] in LDS.
__device__ __forceinline__ void gemm_rows(const unsigned short *A, int lda, const unsigned short *B,
                                          int K, f32x16 acc[4], int lane) {
  for (int nt = 0; nt < 4; nt++)
    for (int r = 0; r < 16; r++) acc[nt][r] = 0.0f;
  const int row = lane & 31, kq = 4 * (lane >> 5);
  for (int ks = 0; ks < K; ks += 8) {
    bf16x4 a;
    for (int j = 0; j < 4; j++) a[j] = (short)A[row * lda + ks + kq + j];
    for (int nt = 0; nt < 4; nt++) {
      bf16x4 bv;
      for (int j = 0; j < 4; j++) bv[j] = (short)B[(ks + kq + j) * kHid + 32 * nt + row];
      acc[nt] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, bv, acc[nt], 0, 0, 0);
    }
  }
}

// Same with B read transposed: B[k][n] = Bt[n][k], Bt row-major [128][K].
__device__ __forceinline__ void gemm_rows_bt(const unsigned short *A, int lda, const unsigned short *Bt,
                                             int K, f32x16 acc[4], int lane) {
  for (int nt = 0; nt < 4; nt++)
    for (int r = 0; r < 16; r++) acc[nt][r] = 0.0f;
  const int row = lane & 31, kq = 4 * (lane >> 5);
  for (int ks = 0; ks < K; ks += 8) {
    bf16x4 a;
    for (int j = 0; j < 4; j++) a[j] = (short)A[row * lda + ks + kq + j];
    for (int nt = 0; nt < 4; nt++) {
      bf16x4 bv;
      for (int j = 0; j < 4; j++) bv[j] = (short)Bt[(32 * nt + row) * K + ks + kq + j];
      acc[nt] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, bv, acc[nt], 0, 0, 0);
    }
  }
}

// Forward of one 32-row tile held in X (LDS, [32][kDin]).  Leaves
// Z1, H1, Z2, H2 (BF16, [32][128]) in LDS and returns y for rows 0..31 in
// lanes 0..31.
__device__ float forward_tile(const unsigned short *X, const unsigned short *W1, const unsigned short *W2,
                              const float *prm, unsigned short *Z1, unsigned short *H1,
                              unsigned short *Z2, unsigned short *H2, int lane) {
  f32x16 acc[4];
  gemm_rows(X, kDin, W1, kDin, acc, lane);
  for (int nt = 0; nt < 4; nt++)
    for (int r = 0; r < 16; r++) {
      const int row = d_row(r, lane), col = 32 * nt + (lane & 31);
      const float z = acc[nt][r] + prm[kB1 + col];
      const unsigned short zb = f2bf(z);
      Z1[row * kHid + col] = zb;
      H1[row * kHid + col] = f2bf(gelu(bf2f(zb)));
    }
  __syncthreads();
  gemm_rows(H1, kHid, W2, kHid, acc, lane);
  for (int nt = 0; nt < 4; nt++)
    for (int r = 0; r < 16; r++) {
      const int row = d_row(r, lane), col = 32 * nt + (lane & 31);
      const float z = acc[nt][r] + prm[kB2 + col];
      const unsigned short zb = f2bf(z);
      Z2[row * kHid + col] = zb;
      H2[row * kHid + col] = f2bf(gelu(bf2f(zb)));
    }
  __syncthreads();
  float y = 0.0f;
  if (lane < 32) {
    for (int cidx = 0; cidx < kHid; cidx++) y += bf2f(H2[lane * kHid + cidx]) * bf2f(f2bf(prm[kW3 + cidx]));
    y += prm[kB3];
  }
  return y;
}

// LDS image of the weights: W1 [kDin][128], W2 [128][128] as BF16.
__device__ void load_weights(unsigned short *W1, unsigned short *W2, const float *prm) {
  for (int i = threadIdx.x; i < kDin * kHid; i += blockDim.x) W1[i] = f2bf(prm[kW1 + i]);
  for (int i = threadIdx.x; i < kHid * kHid; i += blockDim.x) W2[i] = f2bf(prm[kW2 + i]);
}

// Inference on a feature matrix (rows x kDin, BF16): out[row] = MLP(x_row).
// One 64-thread block per 32-row tile.
__global__ void __launch_bounds__(64) k_mlp_forward(size_t rows, const unsigned short *__restrict__ feat,
                                                    const float *__restrict__ prm, float *__restrict__ out) {
  extern __shared__ __align__(16) unsigned char sm_raw[];
  unsigned short *W1 = (unsigned short *)sm_raw;
  unsigned short *W2 = W1 + kDin * kHid;
  unsigned short *X = W2 + kHid * kHid;
  unsigned short *Z1 = X + 32 * kDin, *H1 = Z1 + 32 * kHid, *Z2 = H1 + 32 * kHid, *H2 = Z2 + 32 * kHid;
  const int lane = threadIdx.x;
  load_weights(W1, W2, prm);
  const size_t r0 = (size_t)blockIdx.x * 32;
  for (int i = lane; i < 32 * kDin; i += 64) {
    const size_t row = r0 + i / kDin;
    X[i] = row < rows ? feat[row * kDin + (i % kDin)] : (unsigned short)0;
  }
  __syncthreads();
  const float y = forward_tile(X, W1, W2, prm, Z1, H1, Z2, H2, lane);
  if (lane < 32 && r0 + lane < rows) out[r0 + lane] = y;
}

// Learned preconditioner over all N determinants: C[I] = MLP(features of I).
// Computed transposed, H^T = W^T X^T, with the gfx950 MFMA
// v_mfma_f32_32x32x16_bf16 (K = 16 per instruction: lane l holds A[l%32]
// [8*(l/32) + j] and B[8*(l/32) + j][l%32], j < 8; D as for 32x32x8).  The
// weights are the A operand, read from LDS as one 16-byte load per lane; the
// features the B operand, built in registers from the determinant's strings
// (lane l handles determinant l%32 of a 32-determinant tile).  Layer 1 runs
// 9 k-steps (features 134..143 are zero), 4 accumulators of 32 hidden units.
// The activations never leave the registers: after bias, BF16 rounding,
// GELU and rounding, lane half h of accumulator m holds, in registers
// 8s..8s+7, the hidden units 32m + 16s + 8(j/4) + 4h + j%4 -- which is the B
// operand of one layer-2 k-step when layer 2's k index is permuted the same
// way, so the LDS copy of W2 stores its columns in that order (W2p).  Row
// strides of 19 and 17 16-byte elements (odd) make the A reads free of LDS
// bank conflicts: the 16 lanes of one 16-byte-read lane group
// (2*bit5 + (bit4^bit3^bit2)) have distinct (l%32) mod 16, hence distinct
// element residues.  The weights take 75 KB, so two 8-wave blocks share a
// CU.  Layer 3 reduces each determinant's 128 hidden units over the two
// half-waves.  BF16 rounding points as in forward_tile (X, Z1, H1, Z2, H2,
// w3); accumulation order differs (K = 16 per instruction, the layer-3 sum
// in two halves).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
constexpr int kS1 = 152;           // W1t row stride (BF16): 19 16-byte elements
constexpr int kS2 = 136;           // W2p row stride: 17 16-byte elements
constexpr size_t kPrecondLds = sizeof(unsigned short) * (kHid * kS1 + kHid * kS2) + sizeof(float) * 3 * kHid;

// the hidden unit that layer-2 k position p (16 kk + 8 h + j) carries:
// kk = 2m + s -> 32m + 16s + 8(j/4) + 4h + j%4
__device__ __forceinline__ int w2_unit(int p) {
  const int kk = p >> 4, h = (p >> 3) & 1, j = p & 7;
  return 32 * (kk >> 1) + 16 * (kk & 1) + 8 * (j >> 2) + 4 * h + (j & 3);
}

// WAVES waves per block: 8 (two blocks per CU, registers capped for 4 waves
// per SIMD) or 12 (one block per CU, 3 waves per SIMD, no cap)
template <int WAVES>
__global__ void __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(WAVES == 8 ? 4 : 3)))
k_mlp_precond(size_t N, int nb, const uint64_t *__restrict__ astr,
                                                     const uint64_t *__restrict__ bstr,
                                                     const double *__restrict__ R,
                                                     const double *__restrict__ hii,
                                                     const double *__restrict__ E,
                                                     const float *__restrict__ prm, double *__restrict__ C) {
  extern __shared__ __align__(16) unsigned char sm_raw[];
  unsigned short *W1t = (unsigned short *)sm_raw;  // [kHid][kS1], W1t[h][k] = W1[k][h], zero for k >= 134
  unsigned short *W2p = W1t + kHid * kS1;          // [kHid][kS2], W2p[h][p] = W2[w2_unit(p)][h]
  float *bias1 = (float *)(W2p + kHid * kS2);      // [kHid]
  float *bias2 = bias1 + kHid, *w3 = bias2 + kHid;
  for (int i = threadIdx.x; i < kHid * kS1; i += blockDim.x) {
    const int h = i / kS1, k = i % kS1;  // prm W1 is [kDin][kHid]
    W1t[i] = k < 134 ? f2bf(prm[kW1 + k * kHid + h]) : (unsigned short)0;
  }
  for (int i = threadIdx.x; i < kHid * kS2; i += blockDim.x) {
    const int h = i / kS2, p = i % kS2;  // prm W2 is [kHid (in)][kHid (out)]
    W2p[i] = p < kHid ? f2bf(prm[kW2 + w2_unit(p) * kHid + h]) : (unsigned short)0;
  }
  for (int i = threadIdx.x; i < kHid; i += blockDim.x) {
    bias1[i] = prm[kB1 + i];
    bias2[i] = prm[kB2 + i];
    w3[i] = bf2f(f2bf(prm[kW3 + i]));
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int col = lane & 31, half = lane >> 5;
  const double e = E[0];
  const double sq = sqrt((double)N);
  const uint64_t a0 = astr[0], b0 = bstr[0];
  const float b3 = prm[kB3];
  const size_t ntiles = (N + 31) / 32;
  for (size_t tile = (size_t)blockIdx.x * WAVES + wave; tile < ntiles; tile += (size_t)gridDim.x * WAVES) {
    const size_t I = tile * 32 + col;
    const bool live = I < N;
    const size_t Ic = live ? I : N - 1;
    const size_t ia = Ic / nb, ib = Ic % nb;
    const uint64_t sa = astr[ia], sb = bstr[ib];
    // features 128..135 (half 0) and 136..143 (half 1, zero), as features()
    const double r_scaled = R[Ic] * sq, hmE = hii[Ic] - e;
    double den = -hmE;  // E - h_II
    if (fabs(den) < 1e-3) den = den < 0 ? -1e-3 : 1e-3;
    const double dstep = fmax(-50.0, fmin(50.0, r_scaled / den));
    s16x8 sv = {0, 0, 0, 0, 0, 0, 0, 0};
    if (half == 0) {
      sv[0] = (short)f2bf((float)r_scaled);
      sv[1] = (short)f2bf((float)(hmE / 8.0));
      sv[2] = (short)f2bf((float)(dstep / 10.0));
      sv[3] = (short)f2bf((float)(__popcll(sa ^ a0) / 8.0));
      sv[4] = (short)f2bf((float)(__popcll(sb ^ b0) / 8.0));
      sv[5] = (short)f2bf(1.0f);
    }
    // ---- layer 1: acc[m] = rows 32m..32m+31 of W1^T X^T
    f32x16 acc[4];
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[m][r] = 0.0f;
#pragma unroll 1
    for (int ks = 0; ks < 9; ks++) {
      s16x8 bs;
      if (ks < 8) {
        const unsigned byte = (unsigned)(((ks < 4 ? sa : sb) >> (16 * (ks & 3) + 8 * half)) & 0xffull);
#pragma unroll
        for (int j = 0; j < 8; j++) bs[j] = ((byte >> j) & 1u) ? (short)0x3f80 : (short)0;
      } else {
        bs = sv;
      }
      const bf16x8 bv = __builtin_bit_cast(bf16x8, bs);
#pragma unroll
      for (int m = 0; m < 4; m++)
        acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
            *reinterpret_cast<const bf16x8 *>(W1t + (32 * m + col) * kS1 + 16 * ks + 8 * half), bv, acc[m], 0, 0, 0);
    }
    // ---- H1 = BF16(GELU(BF16(Z1 + b1))) as layer 2's B operands, in registers
    bf16x8 h1[4][2];
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
      for (int s2 = 0; s2 < 2; s2++) {
        s16x8 v;
#pragma unroll
        for (int j = 0; j < 8; j++) {
          const int unit = 32 * m + 16 * s2 + 8 * (j >> 2) + 4 * half + (j & 3);
          const float z = bf2f(f2bf(acc[m][8 * s2 + j] + bias1[unit]));
          v[j] = (short)f2bf(gelu_fast(z));
        }
        h1[m][s2] = __builtin_bit_cast(bf16x8, v);
      }
    // ---- layer 2: acc[m'] = rows 32m'..32m'+31 of W2^T H1^T
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[m][r] = 0.0f;
#pragma unroll
    for (int kk = 0; kk < 8; kk++) {
#pragma unroll
      for (int m = 0; m < 4; m++)
        acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
            *reinterpret_cast<const bf16x8 *>(W2p + (32 * m + col) * kS2 + 16 * kk + 8 * half), h1[kk >> 1][kk & 1],
            acc[m], 0, 0, 0);
      // keep each k-step's four A loads next to their MFMAs (hoisting them
      // all ahead costs more registers than the chip has for this occupancy)
      __builtin_amdgcn_sched_barrier(0);
    }
    // ---- layer 3: y = sum_h w3[h] BF16(GELU(BF16(Z2[h] + b2[h]))) + b3 over both half-waves
    float ysum = 0.f;
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int unit = 32 * m + 8 * (r >> 2) + 4 * half + (r & 3);
        const float z = bf2f(f2bf(acc[m][r] + bias2[unit]));
        ysum += bf2f(f2bf(gelu_fast(z))) * w3[unit];
      }
    ysum += __shfl_xor(ysum, 32, 64);
    if (half == 0 && live) C[I] = (double)(ysum + b3);
  }
}

// Training, stage 1: forward + backward to the activation gradients of one
// 32-sample tile (samples idx[t0..t0+32) of the data set).  Writes, per
// sample: X, H1, H2, dZ1, dZ2 (BF16) and dy (FP32) for the weight-gradient
// stage, and the squared error for the loss.
__global__ void __launch_bounds__(64) k_mlp_backward(size_t nbatch, const uint32_t *__restrict__ idx,
                                                     const unsigned short *__restrict__ feat,
                                                     const float *__restrict__ target,
                                                     const float *__restrict__ prm, float inv_b,
                                                     unsigned short *__restrict__ gX,
                                                     unsigned short *__restrict__ gH1,
                                                     unsigned short *__restrict__ gH2,
                                                     unsigned short *__restrict__ gdZ1,
                                                     unsigned short *__restrict__ gdZ2,
                                                     float *__restrict__ gdy, float *__restrict__ sqerr) {
  extern __shared__ __align__(16) unsigned char sm_raw[];
  unsigned short *W1 = (unsigned short *)sm_raw;
  unsigned short *W2 = W1 + kDin * kHid;
  unsigned short *X = W2 + kHid * kHid;
  unsigned short *Z1 = X + 32 * kDin, *H1 = Z1 + 32 * kHid, *Z2 = H1 + 32 * kHid, *H2 = Z2 + 32 * kHid;
  unsigned short *dZ = H2 + 32 * kHid;
  const int lane = threadIdx.x;
  load_weights(W1, W2, prm);
  const size_t t0 = (size_t)blockIdx.x * 32;
  for (int i = lane; i < 32 * kDin; i += 64) {
    const size_t s = t0 + i / kDin;
    X[i] = s < nbatch ? feat[(size_t)idx[s] * kDin + (i % kDin)] : (unsigned short)0;
  }
  __syncthreads();
  const float y = forward_tile(X, W1, W2, prm, Z1, H1, Z2, H2, lane);
  __shared__ float dys[32];
  if (lane < 32) {
    const size_t s = t0 + lane;
    float dy = 0.0f;
    if (s < nbatch) {
      const float d = y - target[idx[s]];
      dy = 2.0f * d * inv_b;
      sqerr[s] = d * d;
    }
    dys[lane] = dy;
  }
  __syncthreads();
  // dZ2 = dy * w3 * GELU'(Z2)
  for (int i = lane; i < 32 * kHid; i += 64) {
    const int row = i / kHid, col = i % kHid;
    dZ[i] = f2bf(dys[row] * bf2f(f2bf(prm[kW3 + col])) * gelu_d(bf2f(Z2[i])));
  }
  __syncthreads();
  // dH1 = dZ2 * W2^T ; dZ1 = dH1 * GELU'(Z1)
  f32x16 acc[4];
  gemm_rows_bt(dZ, kHid, W2, kHid, acc, lane);
  for (int i = lane; i < 32 * kHid; i += 64) {
    const size_t s = t0 + i / kHid;
    if (s < nbatch) gdZ2[s * kHid + (i % kHid)] = dZ[i];
  }
  __syncthreads();
  for (int nt = 0; nt < 4; nt++)
    for (int r = 0; r < 16; r++) {
      const int row = d_row(r, lane), col = 32 * nt + (lane & 31);
      dZ[row * kHid + col] = f2bf(acc[nt][r] * gelu_d(bf2f(Z1[row * kHid + col])));
    }
  __syncthreads();
  for (int i = lane; i < 32 * kHid; i += 64) {
    const size_t s = t0 + i / kHid;
    if (s < nbatch) {
      gdZ1[s * kHid + (i % kHid)] = dZ[i];
      gH1[s * kHid + (i % kHid)] = H1[i];
      gH2[s * kHid + (i % kHid)] = H2[i];
    }
  }
  for (int i = lane; i < 32 * kDin; i += 64) {
    const size_t s = t0 + i / kDin;
    if (s < nbatch) gX[s * kDin + (i % kDin)] = X[i];
  }
  if (lane < 32 && t0 + lane < nbatch) gdy[t0 + lane] = dys[lane];
}

// Training, stage 2: weight gradients as reductions over the batch on MFMA.
// Block (tile) t of 64 threads computes a 32x32 tile of G = P^T Q where P is
// [nbatch][M] and Q [nbatch][128] (BF16): dW1 = X^T dZ1 (M = 160), dW2 =
// H1^T dZ2 (M = 128).  Samples beyond nbatch are padding (zeros).
__global__ void __launch_bounds__(64) k_mlp_wgrad(size_t nbatch, int M, const unsigned short *__restrict__ P,
                                                  const unsigned short *__restrict__ Q,
                                                  float *__restrict__ G) {
  const int mt = blockIdx.x, nt = blockIdx.y;
  const int lane = threadIdx.x;
  const int mrow = 32 * mt + (lane & 31), ncol = 32 * nt + (lane & 31), kq = 4 * (lane >> 5);
  f32x16 acc;
  for (int r = 0; r < 16; r++) acc[r] = 0.0f;
  for (size_t s0 = 0; s0 < nbatch; s0 += 8) {
    bf16x4 a, b;
    for (int j = 0; j < 4; j++) {
      const size_t s = s0 + kq + j;
      a[j] = s < nbatch ? (short)P[s * M + mrow] : (short)0;
      b[j] = s < nbatch ? (short)Q[s * kHid + ncol] : (short)0;
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
  }
  for (int r = 0; r < 16; r++) G[(size_t)(32 * mt + d_row(r, lane)) * kHid + ncol] = acc[r];
}

// Bias and last-layer gradients: db1 = sum dZ1, db2 = sum dZ2, dw3 = H2^T dy,
// db3 = sum dy (one thread per output, fixed-order sums).
__global__ void k_mlp_smallgrad(size_t nbatch, const unsigned short *__restrict__ dZ1,
                                const unsigned short *__restrict__ dZ2,
                                const unsigned short *__restrict__ H2, const float *__restrict__ dy,
                                float *__restrict__ grad) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < kHid) {
    float s1 = 0.f, s2 = 0.f, s3 = 0.f;
    for (size_t s = 0; s < nbatch; s++) {
      s1 += bf2f(dZ1[s * kHid + c]);
      s2 += bf2f(dZ2[s * kHid + c]);
      s3 += bf2f(H2[s * kHid + c]) * dy[s];
    }
    grad[kB1 + c] = s1;
    grad[kB2 + c] = s2;
    grad[kW3 + c] = s3;
  } else if (c == kHid) {
    float s = 0.f;
    for (size_t i = 0; i < nbatch; i++) s += dy[i];
    grad[kB3] = s;
    grad[kB3 + 1] = grad[kB3 + 2] = grad[kB3 + 3] = 0.f;
  }
}

__global__ void k_adam(int n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ------------------------------------------------------------------ host
static size_t mlp_lds() {
  return sizeof(unsigned short) * ((size_t)kDin * kHid + kHid * kHid + 32 * kDin + 5 * 32 * kHid);
}

static void set_lds(const void *fn) {
  SBDX_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)mlp_lds()));
}

void mlp_ensure(Context &c) {
  if (c.mlp.prm) return;
  c.mlp.prm = (float *)dalloc(c, sizeof(float) * kNParam);
  c.mlp.grad = (float *)dalloc(c, sizeof(float) * kNParam);
  c.mlp.m = (float *)dalloc(c, sizeof(float) * kNParam);
  c.mlp.v = (float *)dalloc(c, sizeof(float) * kNParam);
  SBDX_HIP(hipMemsetAsync(c.mlp.prm, 0, sizeof(float) * kNParam, c.stream));
  SBDX_HIP(hipMemsetAsync(c.mlp.m, 0, sizeof(float) * kNParam, c.stream));
  SBDX_HIP(hipMemsetAsync(c.mlp.v, 0, sizeof(float) * kNParam, c.stream));
}

int mlp_nparam() { return kNParam; }
int mlp_nfeat() { return kDin; }

void mlp_set(Context &c, const float *host) {
  mlp_ensure(c);
  SBDX_HIP(hipMemcpyAsync(c.mlp.prm, host, sizeof(float) * kNParam, hipMemcpyHostToDevice, c.stream));
  SBDX_HIP(hipMemsetAsync(c.mlp.m, 0, sizeof(float) * kNParam, c.stream));
  SBDX_HIP(hipMemsetAsync(c.mlp.v, 0, sizeof(float) * kNParam, c.stream));
  c.mlp.adam_t = 0;
  c.mlp.valid = true;
  SBDX_HIP(hipStreamSynchronize(c.stream));
}

void mlp_get(Context &c, float *host) {
  mlp_ensure(c);
  SBDX_HIP(hipMemcpyAsync(host, c.mlp.prm, sizeof(float) * kNParam, hipMemcpyDeviceToHost, c.stream));
  SBDX_HIP(hipStreamSynchronize(c.stream));
}

void mlp_forward(Context &c, const unsigned short *dfeat, size_t rows, float *dout) {
  set_lds((const void *)k_mlp_forward);
  hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)((rows + 31) / 32)), dim3(64), mlp_lds(), c.stream,
                     rows, dfeat, c.mlp.prm, dout);
  check_launch("k_mlp_forward");
}

// Gradient of the mean squared error over samples idx[0..nbatch) into
// c.mlp.grad; returns the summed squared error.
double mlp_gradient(Context &c, const unsigned short *dfeat, const float *dtarget, const uint32_t *didx,
                    size_t nbatch) {
  const size_t pad = (nbatch + 31) / 32 * 32;
  TmpBuf bX(2 * pad * kDin), bH1(2 * pad * kHid), bH2(2 * pad * kHid), bdZ1(2 * pad * kHid),
      bdZ2(2 * pad * kHid), bdy(4 * pad), bse(4 * pad);
  auto *gX = (unsigned short *)bX.p, *gH1 = (unsigned short *)bH1.p, *gH2 = (unsigned short *)bH2.p;
  auto *gdZ1 = (unsigned short *)bdZ1.p, *gdZ2 = (unsigned short *)bdZ2.p;
  set_lds((const void *)k_mlp_backward);
  hipLaunchKernelGGL(k_mlp_backward, dim3((unsigned)(pad / 32)), dim3(64), mlp_lds(), c.stream, nbatch,
                     didx, dfeat, dtarget, c.mlp.prm, 1.0f / (float)nbatch, gX, gH1, gH2, gdZ1, gdZ2,
                     (float *)bdy.p, (float *)bse.p);
  check_launch("k_mlp_backward");
  hipLaunchKernelGGL(k_mlp_wgrad, dim3(kDin / 32, kHid / 32), dim3(64), 0, c.stream, nbatch, kDin, gX, gdZ1,
                     c.mlp.grad + kW1);
  check_launch("k_mlp_wgrad");
  hipLaunchKernelGGL(k_mlp_wgrad, dim3(kHid / 32, kHid / 32), dim3(64), 0, c.stream, nbatch, kHid, gH1,
                     gdZ2, c.mlp.grad + kW2);
  check_launch("k_mlp_wgrad");
  hipLaunchKernelGGL(k_mlp_smallgrad, dim3(1), dim3(kHid + 1), 0, c.stream, nbatch, gdZ1, gdZ2, gH2,
                     (const float *)bdy.p, c.mlp.grad);
  check_launch("k_mlp_smallgrad");
  std::vector<float> se(nbatch);
  SBDX_HIP(hipMemcpyAsync(se.data(), bse.p, sizeof(float) * nbatch, hipMemcpyDeviceToHost, c.stream));
  SBDX_HIP(hipStreamSynchronize(c.stream));
  double s = 0.0;
  for (float x : se) s += x;
  return s;
}

void mlp_adam_step(Context &c, float lr) {
  c.mlp.adam_t++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)c.mlp.adam_t), c2 = 1.f - std::pow(b2, (float)c.mlp.adam_t);
  hipLaunchKernelGGL(k_adam, dim3((kNParam + 255) / 256), dim3(256), 0, c.stream, kNParam, c.mlp.prm,
                     c.mlp.grad, c.mlp.m, c.mlp.v, lr, b1, b2, 1e-8f, c1, c2);
  check_launch("k_adam");
}

template <int WAVES>
static void launch_precond(Context &c, const double *R, const double *E, double *Cout) {
  auto fn = k_mlp_precond<WAVES>;
  SBDX_HIP(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kPrecondLds));
  // every block loops over its tiles; 4 rounds of resident blocks
  const unsigned g = (unsigned)std::min<size_t>((c.N + 32 * WAVES - 1) / (32 * WAVES), (size_t)c.n_cu * 4);
  hipLaunchKernelGGL(fn, dim3(g), dim3(64 * WAVES), kPrecondLds, c.stream, c.N, (int)c.nb, c.A.str, c.B.str, R,
                     c.hii, E, c.mlp.prm, Cout);
  check_launch("k_mlp_precond");
}

// SBDX_PRECOND_WAVES=8|12 picks the block shape (default 8)
void mlp_precond(Context &c, const double *R, const double *E, double *Cout) {
  static const int waves = [] {
    const char *e = std::getenv("SBDX_PRECOND_WAVES");
    return e && std::string(e) == "12" ? 12 : 8;
  }();
  if (waves == 12)
    launch_precond<12>(c, R, E, Cout);
  else
    launch_precond<8>(c, R, E, Cout);
}

// Features of every determinant for the residual R and Ritz value E (BF16,
// [N][kDin]) -- the inputs the preconditioner sees, for building training
// sets and for tests.
__global__ void k_mlp_features(size_t N, int nb, const uint64_t *__restrict__ astr,
                               const uint64_t *__restrict__ bstr, const double *__restrict__ R,
                               const double *__restrict__ hii, double e, unsigned short *__restrict__ out) {
  const size_t I = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= N) return;
  const size_t ia = I / nb, ib = I % nb;
  features(out + I * kDin, astr[ia], bstr[ib], astr[0], bstr[0], R[I] * sqrt((double)N), hii[I] - e);
}

void mlp_features(Context &c, const double *R, double e, unsigned short *dout) {
  hipLaunchKernelGGL(k_mlp_features, dim3((unsigned)((c.N + 255) / 256)), dim3(256), 0, c.stream, c.N,
                     (int)c.nb, c.A.str, c.B.str, R, c.hii, e, dout);
  check_launch("k_mlp_features");
}

// e = s*Wstar - |d|*Wk  (the error of the step's Ritz vector, oriented toward Wstar)
__global__ void k_err_vec(size_t n, const double *__restrict__ Ws, const double *__restrict__ Wk, double s,
                          double ad, double *__restrict__ e) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    e[i] = s * Ws[i] - ad * Wk[i];
}

__global__ void k_gather_pairs(size_t n, const uint32_t *__restrict__ idx, const unsigned short *__restrict__ feat,
                               const double *__restrict__ e, double scale, unsigned short *__restrict__ fout,
                               float *__restrict__ tout) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * kDin) return;
  const size_t p = t / kDin, k = t % kDin;
  const size_t I = idx[p];
  fout[t] = feat[I * kDin + k];
  if (k == 0) tout[p] = (float)(e[I] * scale);
}

// (feature, target) pairs from the history recorded by the last thick solve:
// target = component of the unit error vector of the step's Ritz vector
// (toward the converged vector in c.W), times sqrt(N).  Up to max_pairs
// pairs, an equal share per recorded step, determinants drawn uniformly.
size_t mlp_collect(Context &c, size_t max_pairs, uint64_t seed, unsigned short *feat_out, float *tgt_out) {
  const int K = c.mlp.rec_n;
  if (K == 0 || max_pairs == 0) return 0;
  const size_t N = c.N;
  const size_t per = std::max<size_t>(1, max_pairs / K);
  TmpBuf fb(sizeof(unsigned short) * N * kDin), eb(sizeof(double) * N), ib(sizeof(uint32_t) * per),
      gf(sizeof(unsigned short) * per * kDin), gt(sizeof(float) * per);
  uint64_t st = seed * 0x9e3779b97f4a7c15ull + 1;
  auto next = [&]() {  // splitmix64
    uint64_t z = (st += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  };
  std::vector<uint32_t> idx(per);
  size_t written = 0;
  for (int k = 0; k < K && written + per <= max_pairs; k++) {
    const double d = device_dot(c, c.W, c.mlp.rec_W[k], N);
    const double s = d >= 0 ? 1.0 : -1.0;
    hipLaunchKernelGGL(k_err_vec, dim3(1024), dim3(256), 0, c.stream, N, c.W, c.mlp.rec_W[k], s, std::fabs(d),
                       (double *)eb.p);
    check_launch("k_err_vec");
    const double en = std::sqrt(device_dot(c, (double *)eb.p, (double *)eb.p, N));
    if (!(en > 1e-12)) continue;
    mlp_features(c, c.mlp.rec_R[k], c.mlp.rec_E[k], (unsigned short *)fb.p);
    for (size_t p = 0; p < per; p++) idx[p] = (uint32_t)(next() % N);
    SBDX_HIP(hipMemcpyAsync(ib.p, idx.data(), sizeof(uint32_t) * per, hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_gather_pairs, dim3((unsigned)((per * kDin + 255) / 256)), dim3(256), 0, c.stream,
                       per, (const uint32_t *)ib.p, (const unsigned short *)fb.p, (const double *)eb.p,
                       std::sqrt((double)N) / en, (unsigned short *)gf.p, (float *)gt.p);
    check_launch("k_gather_pairs");
    SBDX_HIP(hipMemcpyAsync(feat_out + written * kDin, gf.p, sizeof(unsigned short) * per * kDin,
                            hipMemcpyDeviceToHost, c.stream));
    SBDX_HIP(hipMemcpyAsync(tgt_out + written, gt.p, sizeof(float) * per, hipMemcpyDeviceToHost, c.stream));
    SBDX_HIP(hipStreamSynchronize(c.stream));
    written += per;
  }
  return written;
}

// Adam on mini-batches of the uploaded data set; returns the mean squared
// error of each epoch (as seen during the epoch) in loss[epoch].
void mlp_train(Context &c, const unsigned short *dfeat, const float *dtgt, size_t n, int epochs, int batch,
               float lr, uint64_t seed, double *loss) {
  TmpBuf ib(sizeof(uint32_t) * n);
  std::vector<uint32_t> perm(n);
  uint64_t st = seed * 0x9e3779b97f4a7c15ull + 7;
  auto next = [&]() {
    uint64_t z = (st += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  };
  for (size_t i = 0; i < n; i++) perm[i] = (uint32_t)i;
  for (int ep = 0; ep < epochs; ep++) {
    for (size_t i = n; i > 1; i--) std::swap(perm[i - 1], perm[next() % i]);
    SBDX_HIP(hipMemcpyAsync(ib.p, perm.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, c.stream));
    double sse = 0.0;
    for (size_t o = 0; o < n; o += (size_t)batch) {
      const size_t nbt = std::min<size_t>((size_t)batch, n - o);
      sse += mlp_gradient(c, dfeat, dtgt, (const uint32_t *)ib.p + o, nbt);
      mlp_adam_step(c, lr);
    }
    if (loss) loss[ep] = sse / (double)n;
  }
  SBDX_HIP(hipStreamSynchronize(c.stream));
  c.mlp.valid = true;
}

}  // namespace sbdx
