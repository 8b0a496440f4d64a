// sigma = H*C in the alpha-first basis (see common.h), FP64 or FP32.
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
// One wavefront per (row, 64-column chunk), four rows per block.  Workgroups
// are dealt to the 8 XCDs round-robin, so block b runs on XCD b%8; the
// mapping gives XCD g the chunks g, g+8, ... and sweeps every row of one chunk
// before the next, keeping that chunk's slice of x (rows x 64 values)
// resident in the XCD's L2 while all rows reuse it.
template <typename T, bool DIAG>
__global__ void __launch_bounds__(256) k_same_xcd(int nrows, int M, int nch, int nrb,
                                                  const uint64_t *__restrict__ s_off,
                                                  const uint64_t *__restrict__ s_ent,
                                                  const T *__restrict__ s_val,
                                                  const uint64_t *__restrict__ d_off,
                                                  const uint32_t *__restrict__ d_ket,
                                                  const T *__restrict__ d_val,
                                                  const T *__restrict__ dress,
                                                  const T *__restrict__ hii,
                                                  const T *__restrict__ x, T *__restrict__ y) {
  const int g = blockIdx.x & 7;
  const int i = blockIdx.x >> 3;
  const int chunk = g + 8 * (i / nrb);
  const int r = 4 * (i % nrb) + (threadIdx.x >> 6);
  if (chunk >= nch || r >= nrows) return;  // uniform per wavefront
  const int m = chunk * 64 + (threadIdx.x & 63);
  const bool live = m < M;
  const int mm = live ? m : M - 1;  // clamp so every lane reads in bounds
  const size_t Ms = (size_t)M;
  T acc = T(0);
  if (DIAG) acc = hii[r * Ms + mm] * x[r * Ms + mm];
  const uint64_t s0 = s_off[r], s1 = s_off[r + 1];
#pragma unroll 4
  for (uint64_t e = s0; e < s1; e++) {
    const uint64_t ent = s_ent[e];
    const T dr = dress[(size_t)ent_pq(ent) * Ms + mm];
    const T coef = s_val[e] + (ent_neg(ent) ? -dr : dr);
    acc += coef * x[(size_t)ent_ket(ent) * Ms + mm];
  }
  const uint64_t d0 = d_off[r], d1 = d_off[r + 1];
#pragma unroll 8
  for (uint64_t e = d0; e < d1; e++) acc += d_val[e] * x[(size_t)d_ket[e] * Ms + mm];
  if (live) y[r * Ms + m] = acc;
}

// out[c][r] (+)= in[r][c] for an R x Cc row-major matrix, 64x64 LDS tiles.
template <typename T, bool ACC>
__global__ void __launch_bounds__(256) k_transpose(const T *__restrict__ in, T *__restrict__ out,
                                                   int R, int Cc) {
  __shared__ T tile[64][65];
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
      const T v = tile[tx][k];
      const size_t o = (size_t)cc * R + r;
      out[o] = ACC ? out[o] + v : v;
    }
  }
}

// Mixed alpha-beta term, beta singles in sliced-ELL form (SpinTables::ell).
// One block per alpha bra row I.  Each LDS round stages EB alpha singles of
// I: their ket rows x[J,:] and sign-folded ERI rows s_a*(pq|..), with a zero
// at index K for padding entries.  Thread slot = tid + t*1024 walks beta row
// perm[slot]; slices are 64 rows of similar length, so the trip count is
// uniform per wavefront and each entry load is a coalesced 256-byte access
// shared by the EB staged pairs.
template <typename T, int NT, int EB>
__global__ void __launch_bounds__(1024) k_mixed_ell(int nb, int K, int nslots,
                                                    const uint64_t *__restrict__ a_off,
                                                    const uint64_t *__restrict__ a_ent,
                                                    const uint32_t *__restrict__ ell,
                                                    const uint32_t *__restrict__ perm,
                                                    const uint64_t *__restrict__ sl_off,
                                                    const uint32_t *__restrict__ sl_len,
                                                    const T *__restrict__ eri,
                                                    const T *__restrict__ x, T *__restrict__ y) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *lds = reinterpret_cast<T *>(lds_raw);
  const int ldr = nb + K + 1;
  const int I = blockIdx.x;
  const int tid = threadIdx.x;
  T acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) acc[t] = T(0);
  const uint64_t e0 = a_off[I], e1 = a_off[I + 1];
  for (uint64_t e = e0; e < e1; e += EB) {
    const int ne = (int)((e1 - e) < (uint64_t)EB ? (e1 - e) : (uint64_t)EB);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < EB; k++) {
      T *crow = lds + k * ldr;
      T *vrow = crow + nb;
      if (k < ne) {
        const uint64_t ent = a_ent[e + k];
        const size_t J = ent_ket(ent);
        const T sa = ent_neg(ent) ? T(-1) : T(1);
        const T *vsrc = eri + (size_t)ent_pq(ent) * K;
        for (int i = tid; i < nb; i += 1024) crow[i] = x[J * nb + i];
        for (int i = tid; i < K; i += 1024) vrow[i] = sa * vsrc[i];
        if (tid == 0) vrow[K] = T(0);
      } else {
        for (int i = tid; i < ldr; i += 1024) crow[i] = T(0);
      }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const int slot = tid + t * 1024;
      if (slot < nslots) {
        const int sl = slot >> 6;
        const uint32_t len = sl_len[sl];
        const uint32_t *p = ell + sl_off[sl] + (slot & 63);
        T a = T(0);
#pragma unroll 4
        for (uint32_t k = 0; k < len; k++) {
          const uint32_t w = p[64 * (size_t)k];
          const int jb = (int)(w & 0xffffu);
          const int rs = (int)((w >> 16) & 0x7fffu);
          T v = lds[nb + rs] * lds[jb];
#pragma unroll
          for (int q = 1; q < EB; q++) v += lds[q * ldr + nb + rs] * lds[q * ldr + jb];
          a += (w >> 31) ? -v : v;
        }
        acc[t] += a;
      }
    }
  }
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const int slot = tid + t * 1024;
    if (slot < nslots) {
      const uint32_t r = perm[slot];
      if (r != 0xffffffffu) y[(size_t)I * nb + r] += acc[t];
    }
  }
}

// Mixed term when one ket row does not fit in LDS: the row is staged in
// nparts pieces of h kets and each piece is walked with its own sliced-ELL
// table (SpinTables::pell, kets relative to the piece); the sign-folded ERI
// row stays staged for all pieces.
struct PartTables {
  const uint32_t *ell[8];
  const uint64_t *off[8];
  const uint32_t *len[8];
};

template <typename T, int NT>
__global__ void __launch_bounds__(1024) k_mixed_parts(int nb, int K, int nslots, int nparts, int h,
                                                      const uint64_t *__restrict__ a_off,
                                                      const uint64_t *__restrict__ a_ent,
                                                      PartTables pt, const uint32_t *__restrict__ perm,
                                                      const T *__restrict__ eri,
                                                      const T *__restrict__ x, T *__restrict__ y) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *crow = reinterpret_cast<T *>(lds_raw);
  T *vrow = crow + h;
  const int I = blockIdx.x;
  const int tid = threadIdx.x;
  T acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) acc[t] = T(0);
  for (uint64_t e = a_off[I]; e < a_off[I + 1]; e++) {
    const uint64_t ent = a_ent[e];
    const size_t J = ent_ket(ent);
    const T sa = ent_neg(ent) ? T(-1) : T(1);
    const T *vsrc = eri + (size_t)ent_pq(ent) * K;
    __syncthreads();
    for (int i = tid; i < K; i += 1024) vrow[i] = sa * vsrc[i];
    if (tid == 0) vrow[K] = T(0);
    for (int p = 0; p < nparts; p++) {
      const int lo = p * h;
      const int len_p = (nb - lo) < h ? (nb - lo) : h;
      __syncthreads();
      for (int i = tid; i < len_p; i += 1024) crow[i] = x[J * nb + lo + i];
      __syncthreads();
      const uint32_t *ellp = pt.ell[p];
      const uint64_t *offp = pt.off[p];
      const uint32_t *lenp = pt.len[p];
#pragma unroll
      for (int t = 0; t < NT; t++) {
        const int slot = tid + t * 1024;
        if (slot < nslots) {
          const int sl = slot >> 6;
          const uint32_t len = lenp[sl];
          const uint32_t *q = ellp + offp[sl] + (slot & 63);
          T a = T(0);
#pragma unroll 4
          for (uint32_t k = 0; k < len; k++) {
            const uint32_t w = q[64 * (size_t)k];
            const T v = vrow[(w >> 16) & 0x7fffu] * crow[w & 0xffffu];
            a += (w >> 31) ? -v : v;
          }
          acc[t] += a;
        }
      }
    }
  }
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const int slot = tid + t * 1024;
    if (slot < nslots) {
      const uint32_t r = perm[slot];
      if (r != 0xffffffffu) y[(size_t)I * nb + r] += acc[t];
    }
  }
}

// Mixed term with the beta singles in CSR form (used when the packed
// sliced-ELL table cannot hold the indices): x[J,:] and V row in LDS.
template <typename T, int NT>
__global__ void __launch_bounds__(1024) k_mixed_lds(int nb, int K, const uint64_t *__restrict__ a_off,
                                                    const uint64_t *__restrict__ a_ent,
                                                    const uint64_t *__restrict__ b_off,
                                                    const uint64_t *__restrict__ b_ent,
                                                    const T *__restrict__ eri,
                                                    const T *__restrict__ x, T *__restrict__ y) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *crow = reinterpret_cast<T *>(lds_raw);
  T *vrow = crow + nb;
  const int I = blockIdx.x;
  const int tid = threadIdx.x;
  T acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) acc[t] = T(0);
  for (uint64_t e = a_off[I]; e < a_off[I + 1]; e++) {
    const uint64_t ent = a_ent[e];
    const size_t J = ent_ket(ent);
    const T sa = ent_neg(ent) ? T(-1) : T(1);
    const T *vsrc = eri + (size_t)ent_pq(ent) * K;
    __syncthreads();
    for (int i = tid; i < nb; i += 1024) crow[i] = x[J * nb + i];
    for (int i = tid; i < K; i += 1024) vrow[i] = sa * vsrc[i];
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const int ib = tid + t * 1024;
      if (ib < nb) {
        T a = T(0);
        for (uint64_t f = b_off[ib]; f < b_off[ib + 1]; f++) {
          const uint64_t be = b_ent[f];
          const T v = vrow[ent_pq(be)] * crow[ent_ket(be)];
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
template <typename T>
__global__ void __launch_bounds__(256) k_mixed_global(int nb, int nchunk, int K,
                                                      const uint64_t *__restrict__ a_off,
                                                      const uint64_t *__restrict__ a_ent,
                                                      const uint64_t *__restrict__ b_off,
                                                      const uint64_t *__restrict__ b_ent,
                                                      const T *__restrict__ eri,
                                                      const T *__restrict__ x, T *__restrict__ y) {
  const int I = blockIdx.x / nchunk;
  const int ib = (blockIdx.x % nchunk) * blockDim.x + threadIdx.x;
  if (ib >= nb) return;
  const uint64_t f0 = b_off[ib], f1 = b_off[ib + 1];
  T acc = T(0);
  for (uint64_t e = a_off[I]; e < a_off[I + 1]; e++) {
    const uint64_t ent = a_ent[e];
    const T *xr = x + (size_t)ent_ket(ent) * nb;
    const T *vr = eri + (size_t)ent_pq(ent) * K;
    T a = T(0);
    for (uint64_t f = f0; f < f1; f++) {
      const uint64_t be = b_ent[f];
      const T v = vr[ent_pq(be)] * xr[ent_ket(be)];
      a += ent_neg(be) ? -v : v;
    }
    acc += ent_neg(ent) ? -a : a;
  }
  y[(size_t)I * nb + ib] += acc;
}

template <typename From, typename To>
__global__ void __launch_bounds__(256) k_cast(size_t n, const From *__restrict__ in,
                                              To *__restrict__ out) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = (To)in[i];
}

// ------------------------------------------------------------------ host side
template <typename T>
struct SigmaOperands {
  const T *s_val_a, *d_val_a, *s_val_b, *d_val_b, *dress_a, *dress_b, *eri, *hii;
  T *T1, *T2;
};

static SigmaOperands<double> operands(Context &c, double *) {
  return {c.A.s_val, c.A.d_val, c.B.s_val, c.B.d_val, c.A.dress, c.B.dress, c.eri, c.hii, c.T1, c.T2};
}
static SigmaOperands<float> operands(Context &c, float *) {
  return {c.f32.s_val_a, c.f32.d_val_a, c.f32.s_val_b, c.f32.d_val_b, c.f32.dress_a,
          c.f32.dress_b, c.f32.eri,     c.f32.hii,     c.f32.T1,      c.f32.T2};
}

template <typename T>
static void sigma_same(Context &c, const SpinTables &t, const T *s_val, const T *d_val, int nrows,
                       int M, const T *dress, const T *hii, const T *x, T *y, bool diag) {
  const int nch = (int)cdiv(M, 64), nrb = (int)cdiv(nrows, 4);
  const size_t nblk = (size_t)8 * cdiv(nch, 8) * nrb;
  if (nblk >= (1ull << 31)) throw Error(3, "sigma: grid too large for one launch");
  if (diag)
    hipLaunchKernelGGL((k_same_xcd<T, true>), dim3((unsigned)nblk), dim3(256), 0, c.stream, nrows, M,
                       nch, nrb, t.s_off, t.s_ent, s_val, t.d_off, t.d_ket, d_val, dress, hii, x, y);
  else
    hipLaunchKernelGGL((k_same_xcd<T, false>), dim3((unsigned)nblk), dim3(256), 0, c.stream, nrows, M,
                       nch, nrb, t.s_off, t.s_ent, s_val, t.d_off, t.d_ket, d_val, dress, hii, x, y);
  check_launch("k_same_xcd");
}

template <typename T>
static void transpose(Context &c, const T *in, T *out, int R, int Cc, bool acc) {
  const dim3 grid(cdiv(Cc, 64), cdiv(R, 64));
  if (acc)
    hipLaunchKernelGGL((k_transpose<T, true>), grid, dim3(256), 0, c.stream, in, out, R, Cc);
  else
    hipLaunchKernelGGL((k_transpose<T, false>), grid, dim3(256), 0, c.stream, in, out, R, Cc);
  check_launch("k_transpose");
}

void transpose_into(Context &c, const double *in, double *out, int R, int Cc) {
  transpose<double>(c, in, out, R, Cc, false);
}

template <typename T, int NT, int EB>
static void launch_mixed_ell(Context &c, const T *eri, const T *x, T *y, size_t lds) {
  auto fn = k_mixed_ell<T, NT, EB>;
  SBDX_HIP(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(fn, dim3((unsigned)c.na), dim3(1024), lds, c.stream, (int)c.nb, c.K,
                     c.B.nslice * 64, c.A.s_off, c.A.s_ent, c.B.ell, c.B.perm, c.B.sl_off,
                     c.B.sl_len, eri, x, y);
}

template <typename T, int EB>
static bool mixed_ell_dispatch(Context &c, const T *eri, const T *x, T *y) {
  const size_t lds = sizeof(T) * EB * ((size_t)c.nb + c.K + 1);
  const int nt = (c.B.nslice * 64 + 1023) / 1024;
  if (lds > c.lds_limit || nt > 32) return false;
  if (nt <= 1) launch_mixed_ell<T, 1, EB>(c, eri, x, y, lds);
  else if (nt <= 2) launch_mixed_ell<T, 2, EB>(c, eri, x, y, lds);
  else if (nt <= 4) launch_mixed_ell<T, 4, EB>(c, eri, x, y, lds);
  else if (nt <= 8) launch_mixed_ell<T, 8, EB>(c, eri, x, y, lds);
  else if (nt <= 16) launch_mixed_ell<T, 16, EB>(c, eri, x, y, lds);
  else launch_mixed_ell<T, 32, EB>(c, eri, x, y, lds);
  return true;
}

template <typename T, int NT>
static void launch_mixed_lds(Context &c, const T *eri, const T *x, T *y, size_t lds) {
  auto fn = k_mixed_lds<T, NT>;
  SBDX_HIP(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(fn, dim3((unsigned)c.na), dim3(1024), lds, c.stream, (int)c.nb, c.K, c.A.s_off,
                     c.A.s_ent, c.B.s_off, c.B.s_ent, eri, x, y);
}

template <typename T, int NT>
static void launch_mixed_parts(Context &c, const T *eri, const T *x, T *y, size_t lds) {
  auto fn = k_mixed_parts<T, NT>;
  SBDX_HIP(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  PartTables pt{};
  for (int p = 0; p < c.B.nparts; p++) {
    pt.ell[p] = c.B.pell[p];
    pt.off[p] = c.B.psl_off[p];
    pt.len[p] = c.B.psl_len[p];
  }
  hipLaunchKernelGGL(fn, dim3((unsigned)c.na), dim3(1024), lds, c.stream, (int)c.nb, c.K,
                     c.B.nslice * 64, c.B.nparts, c.B.part_h, c.A.s_off, c.A.s_ent, pt, c.B.perm, eri,
                     x, y);
}

// Ket row staged in parts: smallest part count (<= 8) whose piece and ERI row fit.
template <typename T>
static bool mixed_parts_dispatch(Context &c, const T *eri, const T *x, T *y) {
  const int nt = (c.B.nslice * 64 + 1023) / 1024;
  if (nt > 32) return false;
  // Parts already built (e.g. for the other precision) are kept when they fit.
  const int first = c.B.nparts >= 2 ? c.B.nparts : 2;
  for (int P = first; P <= 8; P++) {
    const size_t h = (c.nb + P - 1) / P;
    const size_t lds = sizeof(T) * (h + c.K + 1);
    if (lds > c.lds_limit) continue;
    build_singles_ell_parts(c, c.B, P);
    if (nt <= 1) launch_mixed_parts<T, 1>(c, eri, x, y, lds);
    else if (nt <= 2) launch_mixed_parts<T, 2>(c, eri, x, y, lds);
    else if (nt <= 4) launch_mixed_parts<T, 4>(c, eri, x, y, lds);
    else if (nt <= 8) launch_mixed_parts<T, 8>(c, eri, x, y, lds);
    else if (nt <= 16) launch_mixed_parts<T, 16>(c, eri, x, y, lds);
    else launch_mixed_parts<T, 32>(c, eri, x, y, lds);
    return true;
  }
  return false;
}

template <typename T>
static void sigma_mixed(Context &c, const T *eri, const T *x, T *y) {
  if (c.B.has_ell) {
    // widest staging that fits: 4 pairs (FP32), 2 pairs, then 1, then row parts
    if (sizeof(T) == 4 && mixed_ell_dispatch<T, 4>(c, eri, x, y)) return check_launch("k_mixed_ell");
    if (mixed_ell_dispatch<T, 2>(c, eri, x, y)) return check_launch("k_mixed_ell");
    if (mixed_ell_dispatch<T, 1>(c, eri, x, y)) return check_launch("k_mixed_ell");
    if (mixed_parts_dispatch<T>(c, eri, x, y)) return check_launch("k_mixed_parts");
  }
  const int nb = (int)c.nb;
  const size_t lds = sizeof(T) * ((size_t)nb + c.K);
  const int nt = (nb + 1023) / 1024;
  if (lds <= c.lds_limit && nt <= 32) {
    if (nt <= 1) launch_mixed_lds<T, 1>(c, eri, x, y, lds);
    else if (nt <= 2) launch_mixed_lds<T, 2>(c, eri, x, y, lds);
    else if (nt <= 4) launch_mixed_lds<T, 4>(c, eri, x, y, lds);
    else if (nt <= 8) launch_mixed_lds<T, 8>(c, eri, x, y, lds);
    else if (nt <= 16) launch_mixed_lds<T, 16>(c, eri, x, y, lds);
    else launch_mixed_lds<T, 32>(c, eri, x, y, lds);
  } else {
    const unsigned nchunk = cdiv(nb, 256);
    hipLaunchKernelGGL((k_mixed_global<T>), dim3((unsigned)((size_t)c.na * nchunk)), dim3(256), 0,
                       c.stream, nb, (int)nchunk, c.K, c.A.s_off, c.A.s_ent, c.B.s_off, c.B.s_ent, eri,
                       x, y);
  }
  check_launch("k_mixed");
}

template <typename T>
static void sigma_impl(Context &c, const T *x, T *y) {
  const SigmaOperands<T> o = operands(c, (T *)nullptr);
  const int na = (int)c.na, nb = (int)c.nb;
  const bool f32 = sizeof(T) == 4;
  {
    ProfScope ps(c, f32 ? "same_alpha_f32" : "same_alpha");
    sigma_same<T>(c, c.A, o.s_val_a, o.d_val_a, na, nb, o.dress_b, o.hii, x, y, true);
  }
  {
    ProfScope ps(c, f32 ? "transpose_f32" : "transpose");
    transpose<T>(c, x, o.T1, na, nb, false);
  }
  {
    ProfScope ps(c, f32 ? "same_beta_f32" : "same_beta");
    sigma_same<T>(c, c.B, o.s_val_b, o.d_val_b, nb, na, o.dress_a, nullptr, o.T1, o.T2, false);
  }
  {
    ProfScope ps(c, f32 ? "transpose_add_f32" : "transpose_add");
    transpose<T>(c, o.T2, y, nb, na, true);
  }
  {
    ProfScope ps(c, f32 ? "mixed_f32" : "mixed");
    sigma_mixed<T>(c, o.eri, x, y);
  }
}

void sigma_alpha_first(Context &c, const double *x, double *y) { sigma_impl<double>(c, x, y); }

void sigma_alpha_first_f32(Context &c, const double *x, double *y) {
  ensure_f32(c);
  const unsigned g = (unsigned)std::min<size_t>(cdiv(c.N, 256), 65536);
  hipLaunchKernelGGL((k_cast<double, float>), dim3(g), dim3(256), 0, c.stream, c.N, x, c.f32.x);
  check_launch("k_cast");
  sigma_impl<float>(c, c.f32.x, c.f32.y);
  hipLaunchKernelGGL((k_cast<float, double>), dim3(g), dim3(256), 0, c.stream, c.N, c.f32.y, y);
  check_launch("k_cast");
}

// FP32 copies of every sigma operand, made on first use.
void ensure_f32(Context &c) {
  if (c.f32.ready) return;
  const size_t K = (size_t)c.K;
  auto cast = [&](const double *src, size_t n) {
    float *dst = (float *)dalloc(c, sizeof(float) * (n ? n : 1));
    if (n) {
      const unsigned g = (unsigned)std::min<size_t>(cdiv(n, 256), 65536);
      hipLaunchKernelGGL((k_cast<double, float>), dim3(g), dim3(256), 0, c.stream, n, src, dst);
      check_launch("k_cast");
    }
    return dst;
  };
  c.f32.s_val_a = cast(c.A.s_val, c.A.ns);
  c.f32.d_val_a = cast(c.A.d_val, c.A.nd);
  c.f32.dress_a = cast(c.A.dress, K * c.A.n);
  if (c.B.alias) {
    c.f32.s_val_b = c.f32.s_val_a;
    c.f32.d_val_b = c.f32.d_val_a;
    c.f32.dress_b = c.f32.dress_a;
  } else {
    c.f32.s_val_b = cast(c.B.s_val, c.B.ns);
    c.f32.d_val_b = cast(c.B.d_val, c.B.nd);
    c.f32.dress_b = cast(c.B.dress, K * c.B.n);
  }
  c.f32.eri = cast(c.eri, K * K);
  c.f32.hii = cast(c.hii, c.N);
  c.f32.x = (float *)dalloc(c, sizeof(float) * c.N);
  c.f32.y = (float *)dalloc(c, sizeof(float) * c.N);
  c.f32.T1 = (float *)dalloc(c, sizeof(float) * c.N);
  c.f32.T2 = (float *)dalloc(c, sizeof(float) * c.N);
  SBDX_HIP(hipStreamSynchronize(c.stream));
  c.f32.ready = true;
}

}  // namespace sbdx
