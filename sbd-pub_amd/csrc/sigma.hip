// sigma = H*C in the alpha-first basis (see common.h).
//
// The reference's CPU sigma (chemistry/tpb/mult.h:625-881) walks every
// determinant and evaluates OneExcite/TwoExcite per table entry; its GPU form
// (chemistry/tpb/mult_thrust.h:382-696) launches one thread per (entry x
// other-spin index) with double atomics.  Here the three terms are split by
// spin so every product is a coalesced row operation and no atomics are
// needed:
//   same-spin alpha   y[I,:]  = hii*x + sum_J a_IJ(:) x[J,:]       rows of x
//   same-spin beta    on the transposed vector, same kernel          rows of x^T
//   mixed alpha-beta  y[I,:] += sum_{J in S_a(I)} s_a V_pq * x[J,:] gathered
//                     through the beta-singles table, with x[J,:] and the
//                     ERI row V_pq staged in LDS.
// An alpha single's coefficient depends on the beta string only through
// sum_{k in beta} (pq|kk); that "dressing" is a precomputed [pq][ib] table,
// read coalesced.
#include "common.h"

namespace sbdx {

static inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

constexpr uint64_t kKetMask = 0xffffffffull;
__device__ __forceinline__ uint32_t ent_ket(uint64_t e) { return (uint32_t)(e & kKetMask); }
__device__ __forceinline__ uint32_t ent_pq(uint64_t e) { return (uint32_t)((e >> 32) & 0x7fffffffull); }
__device__ __forceinline__ bool ent_neg(uint64_t e) { return (e >> 63) != 0; }

// Same-spin rows: for bra row r of this spin and every other-spin column m,
//   y[r*M+m] = (DIAG ? hii*x : 0) + sum_singles (v + s*dress[pq][m]) x[J*M+m]
//                                   + sum_doubles v x[J*M+m]
template <bool DIAG>
__global__ void __launch_bounds__(256) k_same(int M, int nchunk, const uint64_t *__restrict__ s_off,
                                              const uint64_t *__restrict__ s_ent,
                                              const double *__restrict__ s_val,
                                              const uint64_t *__restrict__ d_off,
                                              const uint32_t *__restrict__ d_ket,
                                              const double *__restrict__ d_val,
                                              const double *__restrict__ dress,
                                              const double *__restrict__ hii,
                                              const double *__restrict__ x, double *__restrict__ y) {
  const int r = blockIdx.x / nchunk;
  const int m = (blockIdx.x % nchunk) * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const size_t Ms = (size_t)M;
  double acc = 0.0;
  if (DIAG) acc = hii[r * Ms + m] * x[r * Ms + m];
  const uint64_t s0 = s_off[r], s1 = s_off[r + 1];
#pragma unroll 4
  for (uint64_t e = s0; e < s1; e++) {
    const uint64_t ent = s_ent[e];
    const double dr = dress[(size_t)ent_pq(ent) * Ms + m];
    const double coef = s_val[e] + (ent_neg(ent) ? -dr : dr);
    acc += coef * x[(size_t)ent_ket(ent) * Ms + m];
  }
  const uint64_t d0 = d_off[r], d1 = d_off[r + 1];
#pragma unroll 4
  for (uint64_t e = d0; e < d1; e++) acc += d_val[e] * x[(size_t)d_ket[e] * Ms + m];
  y[r * Ms + m] = acc;
}

// out[c][r] (+)= in[r][c] for an R x Cc row-major matrix, 64x64 LDS tiles.
template <bool ACC>
__global__ void __launch_bounds__(256) k_transpose(const double *__restrict__ in,
                                                   double *__restrict__ out, int R, int Cc) {
  __shared__ double tile[64][65];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int c0 = blockIdx.x * 64, r0 = blockIdx.y * 64;
  for (int k = ty; k < 64; k += 4) {
    const int r = r0 + k, cc = c0 + tx;
    if (r < R && cc < Cc) tile[k][tx] = in[(size_t)r * Cc + cc];
  }
  __syncthreads();
  for (int k = ty; k < 64; k += 4) {
    const int cc = c0 + k, r = r0 + tx;
    if (cc < Cc && r < R) {
      const double v = tile[tx][k];
      const size_t o = (size_t)cc * R + r;
      out[o] = ACC ? out[o] + v : v;
    }
  }
}

// Mixed alpha-beta term, one block per alpha bra row I:
//   y[I,ib] += sum_{e in S_a(I)} s_e sum_{f in S_b(ib)} s_f (pq_e|rs_f) x[J_e, jb_f]
// For each alpha single e the ket row x[J_e,:] and s_e*(pq_e|..) are staged
// in LDS; each thread owns NT beta columns and walks their beta singles.
template <int NT>
__global__ void __launch_bounds__(1024) k_mixed_lds(int nb, int K, const uint64_t *__restrict__ a_off,
                                                    const uint64_t *__restrict__ a_ent,
                                                    const uint64_t *__restrict__ b_off,
                                                    const uint64_t *__restrict__ b_ent,
                                                    const double *__restrict__ eri,
                                                    const double *__restrict__ x,
                                                    double *__restrict__ y) {
  extern __shared__ double lds[];
  double *crow = lds;
  double *vrow = lds + nb;
  const int I = blockIdx.x;
  const int tid = threadIdx.x;
  double acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) acc[t] = 0.0;
  const uint64_t e0 = a_off[I], e1 = a_off[I + 1];
  for (uint64_t e = e0; e < e1; e++) {
    const uint64_t ent = a_ent[e];
    const size_t J = ent_ket(ent);
    const double sa = ent_neg(ent) ? -1.0 : 1.0;
    const double *vsrc = eri + (size_t)ent_pq(ent) * K;
    __syncthreads();
    for (int i = tid; i < nb; i += 1024) crow[i] = x[J * nb + i];
    for (int i = tid; i < K; i += 1024) vrow[i] = sa * vsrc[i];
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const int ib = tid + t * 1024;
      if (ib < nb) {
        double a = 0.0;
        const uint64_t f0 = b_off[ib], f1 = b_off[ib + 1];
        for (uint64_t f = f0; f < f1; f++) {
          const uint64_t be = b_ent[f];
          const double v = vrow[ent_pq(be)] * crow[ent_ket(be)];
          a += ent_neg(be) ? -v : v;
        }
        acc[t] += a;
      }
    }
  }
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const int ib = tid + t * 1024;
    if (ib < nb) y[(size_t)I * nb + ib] += acc[t];
  }
}

// Same term without LDS staging (rows too long for LDS): gathers from HBM.
__global__ void __launch_bounds__(256) k_mixed_global(int nb, int nchunk, int K,
                                                      const uint64_t *__restrict__ a_off,
                                                      const uint64_t *__restrict__ a_ent,
                                                      const uint64_t *__restrict__ b_off,
                                                      const uint64_t *__restrict__ b_ent,
                                                      const double *__restrict__ eri,
                                                      const double *__restrict__ x,
                                                      double *__restrict__ y) {
  const int I = blockIdx.x / nchunk;
  const int ib = (blockIdx.x % nchunk) * blockDim.x + threadIdx.x;
  if (ib >= nb) return;
  const uint64_t f0 = b_off[ib], f1 = b_off[ib + 1];
  double acc = 0.0;
  for (uint64_t e = a_off[I]; e < a_off[I + 1]; e++) {
    const uint64_t ent = a_ent[e];
    const double *xr = x + (size_t)ent_ket(ent) * nb;
    const double *vr = eri + (size_t)ent_pq(ent) * K;
    double a = 0.0;
    for (uint64_t f = f0; f < f1; f++) {
      const uint64_t be = b_ent[f];
      const double v = vr[ent_pq(be)] * xr[ent_ket(be)];
      a += ent_neg(be) ? -v : v;
    }
    acc += ent_neg(ent) ? -a : a;
  }
  y[(size_t)I * nb + ib] += acc;
}

template <int NT>
static void launch_mixed_lds(Context &c, const double *x, double *y, size_t lds) {
  auto fn = k_mixed_lds<NT>;
  SBDX_HIP(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(fn, dim3((unsigned)c.na), dim3(1024), lds, c.stream, (int)c.nb, c.K, c.A.s_off,
                     c.A.s_ent, c.B.s_off, c.B.s_ent, c.eri, x, y);
}

static void sigma_same(Context &c, const SpinTables &t, int nrows, int M, const double *dress,
                       const double *hii, const double *x, double *y, bool diag) {
  const unsigned nchunk = cdiv(M, 256);
  const dim3 grid((unsigned)((size_t)nrows * nchunk));
  if (diag)
    hipLaunchKernelGGL(k_same<true>, grid, dim3(256), 0, c.stream, M, (int)nchunk, t.s_off, t.s_ent,
                       t.s_val, t.d_off, t.d_ket, t.d_val, dress, hii, x, y);
  else
    hipLaunchKernelGGL(k_same<false>, grid, dim3(256), 0, c.stream, M, (int)nchunk, t.s_off,
                       t.s_ent, t.s_val, t.d_off, t.d_ket, t.d_val, dress, hii, x, y);
  check_launch("k_same");
}

static void transpose(Context &c, const double *in, double *out, int R, int Cc, bool acc) {
  const dim3 grid(cdiv(Cc, 64), cdiv(R, 64));
  if (acc)
    hipLaunchKernelGGL(k_transpose<true>, grid, dim3(256), 0, c.stream, in, out, R, Cc);
  else
    hipLaunchKernelGGL(k_transpose<false>, grid, dim3(256), 0, c.stream, in, out, R, Cc);
  check_launch("k_transpose");
}

void transpose_into(Context &c, const double *in, double *out, int R, int Cc) {
  transpose(c, in, out, R, Cc, false);
}

void sigma_alpha_first(Context &c, const double *x, double *y) {
  const int na = (int)c.na, nb = (int)c.nb;
  {
    ProfScope ps(c, "same_alpha");
    sigma_same(c, c.A, na, nb, c.B.dress, c.hii, x, y, true);
  }
  {
    ProfScope ps(c, "transpose");
    transpose(c, x, c.T1, na, nb, false);
  }
  {
    ProfScope ps(c, "same_beta");
    sigma_same(c, c.B, nb, na, c.A.dress, nullptr, c.T1, c.T2, false);
  }
  {
    ProfScope ps(c, "transpose_add");
    transpose(c, c.T2, y, nb, na, true);
  }
  {
    ProfScope ps(c, "mixed");
    const size_t lds = sizeof(double) * ((size_t)nb + c.K);
    const int nt = (nb + 1023) / 1024;
    if (lds <= c.lds_limit && nt <= 32) {
      if (nt <= 1) launch_mixed_lds<1>(c, x, y, lds);
      else if (nt <= 2) launch_mixed_lds<2>(c, x, y, lds);
      else if (nt <= 4) launch_mixed_lds<4>(c, x, y, lds);
      else if (nt <= 8) launch_mixed_lds<8>(c, x, y, lds);
      else if (nt <= 16) launch_mixed_lds<16>(c, x, y, lds);
      else launch_mixed_lds<32>(c, x, y, lds);
    } else {
      const unsigned nchunk = cdiv(nb, 256);
      hipLaunchKernelGGL(k_mixed_global, dim3((unsigned)((size_t)na * nchunk)), dim3(256), 0,
                         c.stream, nb, (int)nchunk, c.K, c.A.s_off, c.A.s_ent, c.B.s_off, c.B.s_ent,
                         c.eri, x, y);
    }
    check_launch("k_mixed");
  }
}

}  // namespace sbdx
