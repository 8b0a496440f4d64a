// 1-/2-RDM, chemistry/tpb/correlation.h:378-662 with the element rules of
// chemistry/basic/correlation.h:16-229 (ZeroDiff/OneDiff/TwoDiffCorrelation).
//
// Layout: one[s*L^2 + i + L*j], two[(s+2t)*L^4 + i + L*j + L^2*k + L^3*l].
// Work is done on the alpha-first vector W' = S*W: every contribution of the
// reference, w_bra*w_ket*(interleaved sign), equals w'_bra*w'_ket times the
// one-spin table sign (alpha single, beta single, or their product for the
// mixed term), so the per-pair sums below factor over the other spin:
//   ZeroDiff  -> row/column sums of w'^2 and their occupation-weighted sums
//   OneDiff   -> per single pair: T = s*sum_m w'[I,m]w'[J,m] and
//                Tk[k] = s*sum_m w'[I,m]w'[J,m] n_k(m) (other-spin spectators)
//   TwoDiff   -> per same-spin double pair: T; mixed: T[pq][rs] summed over
//                all (alpha single, beta single) pairs, then placed with the
//                reference's spin-orbital ordering rules.
#include <memory>
#include <mutex>
#include <string>

#include <rocblas/rocblas.h>

#include "common.h"

namespace sbdx {

static inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

__device__ __forceinline__ double parity1r(uint64_t s, int a, int b) {
  int lo = a < b ? a : b, hi = a < b ? b : a;
  uint64_t mask = ((1ull << hi) - 1ull) ^ ((1ull << lo) - 1ull);
  int n = __popcll(s & mask) + (int)((s >> lo) & 1ull);
  return (n & 1) ? -1.0 : 1.0;
}

__device__ __forceinline__ double blk_sum256(double v, double *sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ int row_of(const uint64_t *off, int n, uint64_t e) {
  int lo = 0, hi = n;  // off[lo] <= e < off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= e) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Per row r of w (rows x M): rho[r] = sum_m w^2 and mrow[r][k] = sum_m w^2 n_k(m)
__global__ void __launch_bounds__(256) k_rdm_diag_rows(int M, int L, const uint64_t *__restrict__ ostr,
                                                       const double *__restrict__ w,
                                                       double *__restrict__ rho,
                                                       double *__restrict__ mrow) {
  __shared__ double sh[256];
  __shared__ double acc[64];
  const int r = blockIdx.x;
  if (threadIdx.x < 64) acc[threadIdx.x] = 0.0;
  __syncthreads();
  double s = 0.0;
  for (int m = threadIdx.x; m < M; m += 256) {
    const double v = w[(size_t)r * M + m];
    const double v2 = v * v;
    s += v2;
    if (mrow) {
      uint64_t occ = ostr[m];
      while (occ) {
        const int k = __ffsll((unsigned long long)occ) - 1;
        occ &= occ - 1;
        atomicAdd(&acc[k], v2);
      }
    }
  }
  const double t = blk_sum256(s, sh);
  if (threadIdx.x == 0) rho[r] = t;
  if (mrow && threadIdx.x < L) mrow[(size_t)r * L + threadIdx.x] = acc[threadIdx.x];
}

// ZeroDiff contributions from rho (both spins) and M[p][r] = sum_I n_p(I) mrow[I][r]
__global__ void k_rdm_diag_scatter(int L, int na, int nb, const uint64_t *__restrict__ astr,
                                   const uint64_t *__restrict__ bstr,
                                   const double *__restrict__ rho_a, const double *__restrict__ rho_b,
                                   const double *__restrict__ mrow, double *__restrict__ one,
                                   double *__restrict__ two) {
  // one thread per (p, r) pair of orbitals
  const int p = blockIdx.x, r = threadIdx.x;
  if (r >= L) return;
  const size_t L2 = (size_t)L * L, L4 = L2 * L2;
  double m = 0.0, aa = 0.0, bb = 0.0;
  for (int I = 0; I < na; I++) {
    const uint64_t s = astr[I];
    if ((s >> p) & 1ull) {
      m += mrow[(size_t)I * L + r];
      if (r != p && ((s >> r) & 1ull)) aa += rho_a[I];
    }
  }
  for (int ib = 0; ib < nb; ib++) {
    const uint64_t s = bstr[ib];
    if (((s >> p) & 1ull) && r != p && ((s >> r) & 1ull)) bb += rho_b[ib];
  }
  // alpha-beta pairs: two[2][p,r,p,r] and two[1][r,p,r,p]
  two[2 * L4 + p + L * r + L2 * p + L2 * L * r] += m;
  two[1 * L4 + r + L * p + L2 * r + L2 * L * p] += m;
  // same-spin ordered pairs (p,r), p != r: [p,r,p,r] += w^2 and [p,r,r,p] -= w^2
  if (r != p) {
    two[0 * L4 + p + L * r + L2 * p + L2 * L * r] += aa;
    two[0 * L4 + p + L * r + L2 * r + L2 * L * p] -= aa;
    two[3 * L4 + p + L * r + L2 * p + L2 * L * r] += bb;
    two[3 * L4 + p + L * r + L2 * r + L2 * L * p] -= bb;
  }
  if (r == 0) {
    double da = 0.0, db = 0.0;
    for (int I = 0; I < na; I++)
      if ((astr[I] >> p) & 1ull) da += rho_a[I];
    for (int ib = 0; ib < nb; ib++)
      if ((bstr[ib] >> p) & 1ull) db += rho_b[ib];
    one[p + L * p] += da;
    one[L2 + p + L * p] += db;
  }
}

// Single pairs of one spin: block per pair e; rows of w are this spin's strings.
//   out[e*(L+1)]     = s * sum_m w[I,m] w[J,m]
//   out[e*(L+1)+1+k] = s * sum_m w[I,m] w[J,m] n_k(m)    (other-spin strings)
// Pairs e0 <= e < e1 only; out holds pair e at (e - e0).
__global__ void __launch_bounds__(256) k_rdm_single_pairs(int n, int M, int L, uint64_t e0, uint64_t e1,
                                                          const uint64_t *__restrict__ s_off,
                                                          const uint64_t *__restrict__ s_ent,
                                                          const uint64_t *__restrict__ ostr,
                                                          const double *__restrict__ w,
                                                          double *__restrict__ out) {
  __shared__ double sh[256];
  __shared__ double acc[64];
  for (uint64_t e = e0 + blockIdx.x; e < e1; e += gridDim.x) {
  const int I = row_of(s_off, n, e);
  const uint64_t ent = s_ent[e];
  const size_t J = (uint32_t)(ent & 0xffffffffull);
  const double sg = (ent >> 63) ? -1.0 : 1.0;
  if (threadIdx.x < 64) acc[threadIdx.x] = 0.0;
  __syncthreads();
  double s = 0.0;
  for (int m = threadIdx.x; m < M; m += 256) {
    const double v = w[(size_t)I * M + m] * w[J * M + m];
    s += v;
    if (v != 0.0) {
      uint64_t occ = ostr[m];
      while (occ) {
        const int k = __ffsll((unsigned long long)occ) - 1;
        occ &= occ - 1;
        atomicAdd(&acc[k], v);
      }
    }
  }
  const double t = blk_sum256(s, sh);
  double *o = out + (e - e0) * (size_t)(L + 1);
  if (threadIdx.x == 0) o[0] = sg * t;
  if (threadIdx.x < L) o[1 + threadIdx.x] = sg * acc[threadIdx.x];
  __syncthreads();
  }
}

// Double pairs of one spin, e0 <= e < e1: out[e - e0] = sign * sum_m w[I,m] w[J,m]
__global__ void __launch_bounds__(256) k_rdm_double_pairs(int n, int M, uint64_t e0, uint64_t e1,
                                                          const uint64_t *__restrict__ d_off,
                                                          const uint32_t *__restrict__ d_ket,
                                                          const uint32_t *__restrict__ d_orb,
                                                          const uint64_t *__restrict__ str,
                                                          const double *__restrict__ w,
                                                          double *__restrict__ out) {
  __shared__ double sh[256];
  for (uint64_t e = e0 + blockIdx.x; e < e1; e += gridDim.x) {
  const int I = row_of(d_off, n, e);
  const size_t J = d_ket[e];
  double s = 0.0;
  for (int m = threadIdx.x; m < M; m += 256) s += w[(size_t)I * M + m] * w[J * M + m];
  const double t = blk_sum256(s, sh);
  if (threadIdx.x == 0) {
    const uint32_t o = d_orb[e];
    const int i = o & 255, j = (o >> 8) & 255, a = (o >> 16) & 255, b = (o >> 24) & 255;
    const uint64_t bra = str[I];
    double sg = parity1r(bra, i, a) * parity1r(bra, j, b);
    if (a > j || b < i) sg = -sg;
    out[e - e0] = sg * t;
  }
  }
}

// The same double pairs from the Gram matrix G = w w^T of this spin's n rows
// (one library GEMM per spin instead of a dot product per pair):
// out[e - e0] = sign * G[I, J]
__global__ void __launch_bounds__(256) k_rdm_double_gram(int n, uint64_t e0, uint64_t e1,
                                                         const uint64_t *__restrict__ d_off,
                                                         const uint32_t *__restrict__ d_ket,
                                                         const uint32_t *__restrict__ d_orb,
                                                         const uint64_t *__restrict__ str,
                                                         const double *__restrict__ G,
                                                         double *__restrict__ out) {
  const uint64_t e = e0 + (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= e1) return;
  const int I = row_of(d_off, n, e);
  const size_t J = d_ket[e];
  const uint32_t o = d_orb[e];
  const int i = o & 255, j = (o >> 8) & 255, a = (o >> 16) & 255, b = (o >> 24) & 255;
  const uint64_t bra = str[I];
  double sg = parity1r(bra, i, a) * parity1r(bra, j, b);
  if (a > j || b < i) sg = -sg;
  out[e - e0] = sg * G[(size_t)I * n + J];
}

// OneDiff scatter for the single pairs e0 <= e < e1 of one spin (si =
// spin); T holds pair e at (e - e0).
__global__ void k_rdm_single_scatter(int n, int L, int spin, uint64_t e0, uint64_t e1,
                                     const uint64_t *__restrict__ s_off,
                                     const uint64_t *__restrict__ s_ent,
                                     const uint64_t *__restrict__ str,
                                     const double *__restrict__ T, double *__restrict__ one,
                                     double *__restrict__ two) {
  const uint64_t e = e0 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= e1) return;
  const size_t L2 = (size_t)L * L, L4 = L2 * L2, L3 = L2 * L;
  const int I = row_of(s_off, n, e);
  const uint32_t pq = (uint32_t)((s_ent[e] >> 32) & 0x7fffffffull);
  const int p = pq / L, q = pq % L;  // p = bra-only (cr), q = ket-only (an)
  const double *t = T + (e - e0) * (size_t)(L + 1);
  const double t0 = t[0];
  atomicAdd(&one[spin * L2 + p + (size_t)L * q], t0);
  const size_t bs = (size_t)(3 * spin) * L4;  // same-spin block
  uint64_t occ = str[I];
  while (occ) {
    const int k = __ffsll((unsigned long long)occ) - 1;
    occ &= occ - 1;
    if (k == p) continue;  // the excited orbital's own terms cancel exactly
    atomicAdd(&two[bs + q + L * k + L2 * p + L3 * k], t0);
    atomicAdd(&two[bs + k + L * q + L2 * k + L3 * p], t0);
    atomicAdd(&two[bs + q + L * k + L2 * k + L3 * p], -t0);
    atomicAdd(&two[bs + k + L * q + L2 * p + L3 * k], -t0);
  }
  const size_t b1 = (size_t)(spin + 2 * (1 - spin)) * L4;  // [si + 2 sj]
  const size_t b2 = (size_t)((1 - spin) + 2 * spin) * L4;  // [sj + 2 si]
  for (int k = 0; k < L; k++) {
    const double tk = t[1 + k];
    if (tk == 0.0) continue;
    atomicAdd(&two[b1 + q + L * k + L2 * p + L3 * k], tk);
    atomicAdd(&two[b2 + k + L * q + L2 * k + L3 * p], tk);
  }
}

// TwoDiff scatter for same-spin double pairs e0 <= e < e1 (block 0 alpha,
// 3 beta); T holds pair e at (e - e0).
__global__ void k_rdm_double_scatter(int L, int spin, uint64_t e0, uint64_t e1,
                                     const uint32_t *__restrict__ d_orb,
                                     const double *__restrict__ T, double *__restrict__ two) {
  const uint64_t e = e0 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= e1) return;
  const size_t L2 = (size_t)L * L, L4 = L2 * L2;
  const uint32_t o = d_orb[e];
  const size_t i = o & 255, j = (o >> 8) & 255, a = (o >> 16) & 255, b = (o >> 24) & 255;
  const double t = T[e - e0];
  const size_t bs = (size_t)(3 * spin) * L4;
  atomicAdd(&two[bs + a + L * b + L2 * (i + L * j)], t);
  atomicAdd(&two[bs + b + L * a + L2 * (j + L * i)], t);
  atomicAdd(&two[bs + a + L * b + L2 * (j + L * i)], -t);
  atomicAdd(&two[bs + b + L * a + L2 * (i + L * j)], -t);
}

// Mixed term: Tm[pq*K + rs] += sum_{e in S_a(I)} s_e sum_{ib} w[I,ib] sum_{f in S_b(ib)} s_f w[J_e, jb_f]
template <int NT>
__global__ void __launch_bounds__(1024) k_rdm_mixed(int nb, int K, const uint64_t *__restrict__ a_off,
                                                    const uint64_t *__restrict__ a_ent,
                                                    const uint64_t *__restrict__ b_off,
                                                    const uint64_t *__restrict__ b_ent,
                                                    const double *__restrict__ w,
                                                    double *__restrict__ Tm) {
  extern __shared__ double lds[];
  double *wrow = lds;
  double *trow = lds + nb;
  const int I = blockIdx.x;
  const int tid = threadIdx.x;
  double wi[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const int ib = tid + t * 1024;
    wi[t] = ib < nb ? w[(size_t)I * nb + ib] : 0.0;
  }
  for (uint64_t e = a_off[I]; e < a_off[I + 1]; e++) {
    const uint64_t ent = a_ent[e];
    const size_t J = (uint32_t)(ent & 0xffffffffull);
    const double sa = (ent >> 63) ? -1.0 : 1.0;
    const size_t pq = (ent >> 32) & 0x7fffffffull;
    __syncthreads();
    for (int i = tid; i < nb; i += 1024) wrow[i] = w[J * nb + i];
    for (int i = tid; i < K; i += 1024) trow[i] = 0.0;
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const int ib = tid + t * 1024;
      if (ib < nb && wi[t] != 0.0) {
        for (uint64_t f = b_off[ib]; f < b_off[ib + 1]; f++) {
          const uint64_t be = b_ent[f];
          const double v = wi[t] * wrow[(uint32_t)(be & 0xffffffffull)];
          atomicAdd(&trow[(be >> 32) & 0x7fffffffull], (be >> 63) ? -v : v);
        }
      }
    }
    __syncthreads();
    for (int i = tid; i < K; i += 1024) {
      const double v = trow[i];
      if (v != 0.0) atomicAdd(&Tm[pq * K + i], sa * v);
    }
  }
}

// The same mixed term with no atomics: one block per alpha pair index pq
// loops over the alpha singles (I -> J, s_a) with that pq (a_pm), staging
// x[I,:] and x[J,:] in LDS (each followed by a zero).  The beta singles are
// grouped by their pair index rs into bins laid out as a sliced ELL (b_pb:
// 64 bins per slice, entry k of lane l at b_off[s] + 64 k + l, packed
// ib | jb << 15 | (s_b<0) << 31, padding reading the zeros): wavefront w
// owns slices w, w + 16, ..., each lane one bin, and adds s_a times
// sum s_b x[I,ib] x[J,jb] over its bin to a register; after the last alpha
// single the registers are Tm's row pq.  Table loads are coalesced (one
// 256-byte row per wavefront step) and no two lanes share a bin.  Needs
// nb < 8192 (the staged rows: at most 8 values per thread) and at most
// 16 * MS slices.
template <int MS>
__global__ void __launch_bounds__(1024) k_rdm_mixed_pq(int nb, int K, int nslice,
                                                       const uint64_t *__restrict__ a_off,
                                                       const uint64_t *__restrict__ a_ent,
                                                       const uint32_t *__restrict__ b_pb,
                                                       const uint64_t *__restrict__ b_off,
                                                       const uint32_t *__restrict__ b_len,
                                                       const uint32_t *__restrict__ b_perm,
                                                       const double *__restrict__ w, double *__restrict__ Tm) {
  extern __shared__ double lds[];
  double *xI = lds, *xJ = lds + nb + 1;
  const int pq = blockIdx.x;
  const uint64_t e0 = a_off[pq], e1 = a_off[pq + 1];
  if (e0 == e1) return;  // uniform per block, before any barrier (Tm's row stays 0)
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  double acc[MS];
#pragma unroll
  for (int j = 0; j < MS; j++) acc[j] = 0.0;
  for (uint64_t e = e0; e < e1; e++) {
    const uint64_t ent = a_ent[e];
    const size_t I = (uint32_t)(ent & 0xffffffffull);
    const size_t J = (uint32_t)((ent >> 32) & 0x7fffffffull);
    const double sa = (ent >> 63) ? -1.0 : 1.0;
    // both rows loaded before the barrier that ends the previous alpha
    // single's reads (nb + 1 <= 8192: at most 8 values per thread and row)
    double vi[8], vj[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int i = tid + k * 1024;
      vi[k] = i < nb ? w[I * nb + i] : 0.0;
      vj[k] = i < nb ? w[J * nb + i] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int i = tid + k * 1024;
      if (i <= nb) {
        xI[i] = vi[k];
        xJ[i] = vj[k];
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < MS; j++) {
      const int sl = wv + 16 * j;
      if (sl < nslice) {
        const uint32_t len = b_len[sl];
        const uint32_t *p = b_pb + b_off[sl] + lane;
        double s = 0.0;
#pragma unroll 4
        for (uint32_t k = 0; k < len; k++) {
          const uint32_t wd = p[64 * (size_t)k];
          const double v = xI[wd & 0x7fffu] * xJ[(wd >> 15) & 0x7fffu];
          s += (wd >> 31) ? -v : v;
        }
        acc[j] += sa * s;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < MS; j++) {
    const int sl = wv + 16 * j;
    if (sl < nslice) {
      const uint32_t bin = b_perm[sl * 64 + lane];
      if (bin != 0xffffffffu) Tm[(size_t)pq * K + bin] = acc[j];
    }
  }
}

// Place Tm[pq][rs] (alpha p->q, beta r->s) with TwoDiffCorrelation's rules on
// the sorted spin orbitals I<J (bra-only 2p, 2r+1) and A<B (ket-only 2q, 2s+1).
__global__ void k_rdm_mixed_place(int L, const double *__restrict__ Tm, double *__restrict__ two) {
  const size_t K = (size_t)L * L;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= K * K) return;
  const double t = Tm[idx];
  if (t == 0.0) return;
  const int pq = (int)(idx / K), rs = (int)(idx % K);
  const int p = pq / L, q = pq % L, r = rs / L, s = rs % L;
  const int i = 2 * p, j = 2 * r + 1, a = 2 * q, b = 2 * s + 1;
  const int I = i < j ? i : j, J = i < j ? j : i, A = a < b ? a : b, B = a < b ? b : a;
  const size_t oi = I >> 1, si = I & 1, oa = A >> 1, sa = A & 1;
  const size_t oj = J >> 1, sj = J & 1, ob = B >> 1;
  const size_t L2 = K, L4 = K * K;
  // distinct (pq, rs) can land on one element, hence the atomics
  if (si == sa) {
    atomicAdd(&two[(si + 2 * sj) * L4 + oa + L * ob + L2 * (oi + L * oj)], t);
    atomicAdd(&two[(sj + 2 * si) * L4 + ob + L * oa + L2 * (oj + L * oi)], t);
  } else {
    atomicAdd(&two[(si + 2 * sj) * L4 + oa + L * ob + L2 * (oj + L * oi)], t);
    atomicAdd(&two[(sj + 2 * si) * L4 + ob + L * oa + L2 * (oi + L * oj)], t);
  }
}

void transpose_into(Context &c, const double *in, double *out, int R, int Cc);

template <int NT>
static void launch_rdm_mixed(Context &c, const double *w, double *Tm, size_t lds) {
  auto fn = k_rdm_mixed<NT>;
  SBDX_HIP(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(fn, dim3((unsigned)c.na), dim3(1024), lds, c.stream, (int)c.nb, c.K, c.A.s_off,
                     c.A.s_ent, c.B.s_off, c.B.s_ent, w, Tm);
  check_launch("k_rdm_mixed");
}

// p[0, n) = 0 (a fill kernel rather than hipMemsetAsync, so the work is a
// plain kernel sequence in every path)
__global__ void __launch_bounds__(256) k_rdm_zero(size_t n, double *__restrict__ p) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = 0.0;
}

static void zero(Context &c, double *p, size_t n) {
  hipLaunchKernelGGL(k_rdm_zero, dim3((unsigned)std::min<size_t>(cdiv(n, 256), 4096)), dim3(256), 0, c.stream,
                     n, p);
  check_launch("k_rdm_zero");
}

// Working memory of one RDM evaluation: per-string sums, the mixed-term
// accumulator Tm[K][K], and room for ts_cap single pairs ((L+1) values each)
// and td_cap double pairs (the pair tables are processed in chunks of that
// many entries).
struct RdmWork {
  double *rho_a, *rho_b, *mrow, *Tm, *Ts, *Td;
  size_t ts_cap, td_cap;
  double *G = nullptr;  // room for max(na, nb)^2 values, or null (dot products per double pair)
};

// Everything after the sign conversion: W' (alpha-first, device) to one and
// two (device, overwritten).  Kernels only, on c.stream; synchronizes once.
static void rdm_run(Context &c, const double *Wp, double *one, double *two, const RdmWork &wk) {
  const int L = c.L;
  const size_t L2 = (size_t)L * L, L4 = L2 * L2, K = (size_t)c.K;
  const int na = (int)c.na, nb = (int)c.nb;
  const size_t lds = sizeof(double) * ((size_t)nb + K);
  const int nt = (nb + 1023) / 1024;
  // the pair-major kernel stages two rows, each followed by a zero
  const size_t lds_pq = sizeof(double) * (2 * (size_t)nb + 2);
  const bool by_pair = c.A.pm_ent && c.B.pb && c.B.pb_nslice <= 64 && nb < 8192 && lds_pq <= c.lds_limit;
  if (!by_pair && (lds > c.lds_limit || nt > 32))
    throw Error(3, "rdm: beta string count too large for the LDS-staged mixed RDM kernel");
  zero(c, one, 2 * L2);
  zero(c, two, 4 * L4);
  zero(c, wk.Tm, K * K);
  // W'^T (beta-major) for the beta-side sums
  transpose_into(c, Wp, c.T1, na, nb);
  // ZeroDiff
  hipLaunchKernelGGL(k_rdm_diag_rows, dim3(na), dim3(256), 0, c.stream, nb, L, c.B.str, Wp, wk.rho_a, wk.mrow);
  check_launch("k_rdm_diag_rows");
  hipLaunchKernelGGL(k_rdm_diag_rows, dim3(nb), dim3(256), 0, c.stream, na, L, c.A.str, c.T1, wk.rho_b,
                     (double *)nullptr);
  check_launch("k_rdm_diag_rows");
  hipLaunchKernelGGL(k_rdm_diag_scatter, dim3(L), dim3(64), 0, c.stream, L, na, nb, c.A.str, c.B.str,
                     wk.rho_a, wk.rho_b, wk.mrow, one, two);
  check_launch("k_rdm_diag_scatter");
  // same-spin singles and doubles, alpha (rows of W') then beta (rows of
  // W'^T), in chunks of the pair tables
  for (int spin = 0; spin < 2; spin++) {
    const SpinTables &t = spin ? c.B : c.A;
    const SpinTables &o = spin ? c.A : c.B;
    const double *w = spin ? c.T1 : Wp;
    const int M = spin ? na : nb;
    for (size_t e0 = 0; e0 < t.ns; e0 += wk.ts_cap) {
      const size_t e1 = std::min(t.ns, e0 + wk.ts_cap);
      hipLaunchKernelGGL(k_rdm_single_pairs, dim3((unsigned)std::min<size_t>(e1 - e0, 1u << 20)), dim3(256), 0,
                         c.stream, t.n, M, L, (uint64_t)e0, (uint64_t)e1, t.s_off, t.s_ent, o.str, w, wk.Ts);
      check_launch("k_rdm_single_pairs");
      hipLaunchKernelGGL(k_rdm_single_scatter, dim3(cdiv(e1 - e0, 256)), dim3(256), 0, c.stream, t.n, L, spin,
                         (uint64_t)e0, (uint64_t)e1, t.s_off, t.s_ent, t.str, wk.Ts, one, two);
      check_launch("k_rdm_single_scatter");
    }
    const bool gram = wk.G && t.nd > 0;
    if (gram) {
      // G = w w^T: w is n rows of M values, i.e. the column-major M x n
      // matrix B, and G = B^T B (n x n, symmetric)
      const double one = 1.0, zero_d = 0.0;
      if (rocblas_dgemm((rocblas_handle)blas_handle(c), rocblas_operation_transpose, rocblas_operation_none, t.n,
                        t.n, M, &one, w, M, w, M, &zero_d, wk.G, t.n) != rocblas_status_success)
        throw Error(3, "rdm: rocblas_dgemm (Gram matrix) failed");
    }
    for (size_t e0 = 0; e0 < t.nd; e0 += wk.td_cap) {
      const size_t e1 = std::min(t.nd, e0 + wk.td_cap);
      if (gram) {
        hipLaunchKernelGGL(k_rdm_double_gram, dim3((unsigned)cdiv(e1 - e0, 256)), dim3(256), 0, c.stream, t.n,
                           (uint64_t)e0, (uint64_t)e1, t.d_off, t.d_ket, t.d_orb, t.str, wk.G, wk.Td);
        check_launch("k_rdm_double_gram");
      } else {
        hipLaunchKernelGGL(k_rdm_double_pairs, dim3((unsigned)std::min<size_t>(e1 - e0, 1u << 20)), dim3(256),
                           0, c.stream, t.n, M, (uint64_t)e0, (uint64_t)e1, t.d_off, t.d_ket, t.d_orb, t.str, w,
                           wk.Td);
        check_launch("k_rdm_double_pairs");
      }
      hipLaunchKernelGGL(k_rdm_double_scatter, dim3(cdiv(e1 - e0, 256)), dim3(256), 0, c.stream, L, spin,
                         (uint64_t)e0, (uint64_t)e1, t.d_orb, wk.Td, two);
      check_launch("k_rdm_double_scatter");
    }
  }
  // mixed alpha-single x beta-single
  if (by_pair) {
    const int ms = (c.B.pb_nslice + 15) / 16;
    auto go = [&](auto fn) {
      SBDX_HIP(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_pq));
      hipLaunchKernelGGL(fn, dim3((unsigned)K), dim3(1024), lds_pq, c.stream, nb, (int)K, c.B.pb_nslice,
                         c.A.pm_off, c.A.pm_ent, c.B.pb, c.B.pb_off, c.B.pb_len, c.B.pb_perm, Wp, wk.Tm);
    };
    if (ms <= 1) go(k_rdm_mixed_pq<1>);
    else if (ms <= 2) go(k_rdm_mixed_pq<2>);
    else go(k_rdm_mixed_pq<4>);
    check_launch("k_rdm_mixed_pq");
  } else if (nt <= 1) launch_rdm_mixed<1>(c, Wp, wk.Tm, lds);
  else if (nt <= 2) launch_rdm_mixed<2>(c, Wp, wk.Tm, lds);
  else if (nt <= 4) launch_rdm_mixed<4>(c, Wp, wk.Tm, lds);
  else if (nt <= 8) launch_rdm_mixed<8>(c, Wp, wk.Tm, lds);
  else if (nt <= 16) launch_rdm_mixed<16>(c, Wp, wk.Tm, lds);
  else launch_rdm_mixed<32>(c, Wp, wk.Tm, lds);
  hipLaunchKernelGGL(k_rdm_mixed_place, dim3(cdiv(K * K, 256)), dim3(256), 0, c.stream, L, wk.Tm, two);
  check_launch("k_rdm_mixed_place");
  SBDX_HIP(hipStreamSynchronize(c.stream));
}

void compute_rdm(Context &c, const double *Wp, double *one, double *two) {
  ProfScope ps(c, "rdm");
  const size_t L = (size_t)c.L, K = (size_t)c.K;
  const size_t ns = std::max<size_t>(std::max(c.A.ns, c.B.ns), 1);
  const size_t nd = std::max<size_t>(std::max(c.A.nd, c.B.nd), 1);
  TmpBuf b_rho_a(sizeof(double) * c.na), b_rho_b(sizeof(double) * c.nb), b_mrow(sizeof(double) * c.na * L),
      b_tm(sizeof(double) * K * K), b_ts(sizeof(double) * ns * (L + 1)), b_td(sizeof(double) * nd);
  RdmWork wk{(double *)b_rho_a.p, (double *)b_rho_b.p, (double *)b_mrow.p, (double *)b_tm.p,
             (double *)b_ts.p,    (double *)b_td.p,    ns,                  nd};
  // the Gram matrices of the double pairs when a quarter of free memory
  // holds one (else a dot product per pair)
  const size_t mx = std::max(c.na, c.nb);
  size_t free_b = 0, total_b = 0;
  SBDX_HIP(hipMemGetInfo(&free_b, &total_b));
  std::unique_ptr<TmpBuf> b_g;
  if (mx < (1u << 31) && sizeof(double) * mx * mx < free_b / 4) {
    b_g.reset(new TmpBuf(sizeof(double) * mx * mx));
    wk.G = (double *)b_g->p;
  }
  rdm_run(c, Wp, one, two, wk);
}

// ---------------------------------------------------------------------------
// The same evaluation with working memory from a fixed static device array,
// for callers that cannot allocate (sbdx_rdm_device_noalloc).  Each device
// holds its own copy of `scratch` (kRdmScratchDoubles doubles = 128 MiB) from
// the first use of this file's kernels until the process ends.  In host code
// the array's name is only a stand-in: its device address comes from
// hipGetSymbolAddress, once per device, and only that address is passed on.
// Every call on one device shares the array, so calls that overlap on one
// device (other threads, other handles) race; the array keeps the last
// call's contents, and rdm_run clears what must start at zero.
constexpr size_t kRdmScratchDoubles = size_t(16) << 20;
static __device__ __align__(256) double scratch[kRdmScratchDoubles];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kMaxDevices] = {};

static double *scratch_address(int device) {
  if (device < 0 || device >= kMaxDevices) throw Error(3, "rdm (no-alloc): device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = -1;
    SBDX_HIP(hipGetDevice(&cur));
    if (cur != device) throw Error(1, "rdm (no-alloc): call with the handle's device current");
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw Error(2, std::string("rdm (no-alloc): hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_ptr[device] = (double *)p;
  }
  return g_scratch_ptr[device];
}

size_t rdm_noalloc_fixed_doubles(const Context &c) {
  auto up = [](size_t n) { return (n + 31) & ~size_t(31); };  // 256-byte aligned regions
  return up(c.na) + up(c.nb) + up(c.na * (size_t)c.L) + up((size_t)c.K * c.K);
}

void compute_rdm_noalloc(Context &c, const double *Wp, double *one, double *two) {
  auto up = [](size_t n) { return (n + 31) & ~size_t(31); };
  const size_t L = (size_t)c.L, K = (size_t)c.K;
  const size_t fixed = rdm_noalloc_fixed_doubles(c);
  // at least 64 single pairs and 64 double pairs per chunk
  if (fixed + up(64 * (L + 1)) + up(64) > kRdmScratchDoubles) {
    char msg[256];
    snprintf(msg, sizeof msg,
             "rdm (no-alloc): the instance needs %zu + %zu doubles of working memory, more than the %zu "
             "(128 MiB) of the static scratch array",
             fixed, up(64 * (L + 1)) + up(64), kRdmScratchDoubles);
    throw Error(1, msg);
  }
  const int device = c.device;
  double *base = scratch_address(device);
  double *p = base;
  RdmWork wk{};
  wk.rho_a = p, p += up(c.na);
  wk.rho_b = p, p += up(c.nb);
  wk.mrow = p, p += up(c.na * L);
  wk.Tm = p, p += up(K * K);
  const size_t rem = kRdmScratchDoubles - (size_t)(p - base);
  const size_t ns = std::max<size_t>(std::max(c.A.ns, c.B.ns), 1);
  const size_t nd = std::max<size_t>(std::max(c.A.nd, c.B.nd), 1);
  // half the rest for single pairs, the remainder for double pairs
  wk.ts_cap = std::min(ns, std::max<size_t>(64, (rem / 2) / (L + 1)));
  wk.Ts = p, p += up(wk.ts_cap * (L + 1));
  wk.td_cap = std::min(nd, kRdmScratchDoubles - (size_t)(p - base));
  wk.Td = p;
  rdm_run(c, Wp, one, two, wk);
}

}  // namespace sbdx
