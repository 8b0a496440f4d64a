// sigma = H*C in the alpha-first basis (see common.h), FP64 or FP32.
//
// The reference's CPU sigma (chemistry/tpb/mult.h:625-881) walks every
// determinant and evaluates OneExcite/TwoExcite per table entry; its GPU form
// (chemistry/tpb/mult_thrust.h:382-696) launches one thread per (entry x
// other-spin index) with double atomics.  Here the three terms are split by
// spin so every product is a coalesced row operation and no atomics are
// needed:
//   same-spin         y = hii*x + H_a X + X H_b with X the na x nb vector:
//                     either two dense GEMMs (the same-spin blocks H_a, H_b
//                     are symmetric and, at selected-CI sizes, ~5-15% full
//                     with no empty tiles) or sparse row kernels (x and its
//                     transpose); the faster is picked on first use
//   mixed alpha-beta  y[I,:] += sum_{J in S_a(I)} s_a V_pq * x[J,:] gathered
//                     through the beta-singles table, with x[J,:] and the
//                     ERI row V_pq staged in LDS.
// A single's matrix element depends on the other spin's string through
// sum_{k in other} (pq|kk) (the "dressing"); that part is evaluated in the
// mixed kernels from the staged ERI row's diagonal (alpha singles) and from
// one extra staged row per bra, sum_{k in alpha(I)} (kk|..), with x[I,:]
// (beta singles), so the same-spin blocks are plain matrices.
#include <rocblas/rocblas.h>

#include <algorithm>
#include <array>
#include <cstdlib>
#include <string>
#include <vector>

#include "common.h"

namespace sbdx {

static inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

constexpr uint64_t kKetMask = 0xffffffffull;
__device__ __forceinline__ uint32_t ent_ket(uint64_t e) { return (uint32_t)(e & kKetMask); }
__device__ __forceinline__ uint32_t ent_pq(uint64_t e) { return (uint32_t)((e >> 32) & 0x7fffffffull); }
__device__ __forceinline__ bool ent_neg(uint64_t e) { return (e >> 63) != 0; }

// sum over the occupied orbitals k of `occ` of v[k*(L+1)], the (kk) entries
// of an ERI row; every string of a spin has the same count, so the trip
// count is uniform across a wavefront.
template <typename T>
__device__ __forceinline__ T diag_sum(const T *v, uint64_t occ, int L1) {
  T s = T(0);
  while (occ) {
    s += v[__builtin_ctzll(occ) * L1];
    occ &= occ - 1;
  }
  return s;
}

// The spectator row of bra I: w[i] = sum_{k in alpha(I)} (kk|i) for the
// off-diagonal ERI indices i (zero on the (rr) diagonal, which no beta single
// addresses), so the dressing kernels pick up beta singles' alpha dressing.
template <typename T>
__device__ __forceinline__ T spectator_val(const T *__restrict__ eri, uint64_t occ, int L1, int K,
                                           int i) {
  if (i % L1 == 0) return T(0);
  T s = T(0);
  while (occ) {
    s += eri[(size_t)(__builtin_ctzll(occ) * L1) * K + i];
    occ &= occ - 1;
  }
  return s;
}

// Same-spin rows: for bra row r of this spin and every other-spin column m,
//   y[r*M+m] = (DIAG ? hii*x : 0) + sum_singles v x[J*M+m] + sum_doubles v x[J*M+m]
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
  for (uint64_t e = s0; e < s1; e++) acc += s_val[e] * x[(size_t)ent_ket(s_ent[e]) * Ms + mm];
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

// EB values of one element type read as one vector (one LDS instruction of
// 4, 8 or 16 bytes); EB == 1 is the scalar.
template <typename T, int EB> struct Pack { typedef T V __attribute__((ext_vector_type(EB))); };
template <typename T> struct Pack<T, 1> { typedef T V; };
template <typename T, int EB> __device__ __forceinline__ T &lane_of(typename Pack<T, EB>::V &v, int q) {
  return reinterpret_cast<T *>(&v)[q];
}
template <typename T, int EB> __device__ __forceinline__ T lane_c(const typename Pack<T, EB>::V &v, int q) {
  return reinterpret_cast<const T *>(&v)[q];
}
// v with the sign bit sb (0 or 0x80000000) flipped in each of its EB values
// (high word of a double): one XOR per value, so the product folds into an FMA.
__device__ __forceinline__ float flip_sign(float v, uint32_t sb) {
  return __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, v) ^ sb);
}
__device__ __forceinline__ double flip_sign(double v, uint32_t sb) {
  return __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, v) ^ ((uint64_t)sb << 32));
}
template <typename T, int EB>
__device__ __forceinline__ typename Pack<T, EB>::V flip_sign(typename Pack<T, EB>::V v, uint32_t sb) {
#pragma unroll
  for (int q = 0; q < EB; q++) lane_of<T, EB>(v, q) = flip_sign(lane_of<T, EB>(v, q), sb);
  return v;
}
template <typename T, int EB> __device__ __forceinline__ T hsum(const typename Pack<T, EB>::V &v) {
  T s = reinterpret_cast<const T *>(&v)[0];
#pragma unroll
  for (int q = 1; q < EB; q++) s += reinterpret_cast<const T *>(&v)[q];
  return s;
}

// Mixed alpha-beta term, beta singles in the packed sliced-ELL form
// (SpinTables::ell2).  One block per alpha bra row I.  Each LDS round stages
// EB "pairs": for an alpha single of I its sign-folded ERI row s_a*(pq|..)
// and ket row x[J,:]; after the singles, one spectator pair with the row
// sum_{k in alpha(I)} (kk|..) (spectator_val) and x[I,:].  The pairs are
// interleaved, W[i][q] = v_q[i] at LDS elements 0..K-1 and X[j][q] = x_q[j]
// at elements K..K+nb-1, then one zero element (padding entries), so a
// table word holds both byte offsets pre-shifted (entry = (K + ket) << 4 |
// pq << 19 | sign << 31 for 16-byte elements) and one entry costs one mask,
// one shift-and-mask, two LDS reads of EB values (4 to 32 bytes: gfx950
// serves a 16-byte read per quarter-wavefront at about the cycles of a
// 4-byte one), the EB-term dot product and one signed add.  Thread slot =
// tid + t*1024 walks beta row perm[slot]; slices are 64 rows of similar
// length, so the trip count is uniform per wavefront and each entry load is
// a coalesced 256-byte access shared by the EB staged pairs.  Per pair the
// slot also adds the alpha single's dressing x[J,r] * sum_{k in beta(r)} s_a
// (pq|kk) from the staged ERI row's diagonal (zero in the spectator row).
// A row's slots may be split over gridDim.y blocks of 1024*NT slots each
// (fewer registers per thread, more resident waves, the row staged once per
// block).
// SYM (equal string sets, x = x^T or, odd != 0, x = -x^T): block p handles
// bra I = perm[p] and only the slots s >= p of the length-sorted table,
// adding each value at y[I,r] and, for s > p, +-value at y[r,I] (r =
// perm[s]): every unordered pair {I, r} is produced once, and element (a,b)
// is written by exactly one block (a's when pos(b) >= pos(a), b's
// otherwise); the odd sector leaves the diagonal 0.  Blocks start at the
// slice of position p and exit when their chunk lies past the table; p
// ascending puts the longest rows first.
template <typename T, int NT, int EB, bool SYM>
__global__ void __launch_bounds__(1024) k_mixed_ell(int nb, int K, int L, int nslots,
                                                    const uint64_t *__restrict__ a_off,
                                                    const uint64_t *__restrict__ a_ent,
                                                    const uint64_t *__restrict__ a_str,
                                                    const uint64_t *__restrict__ b_str,
                                                    const uint32_t *__restrict__ ell,
                                                    const uint32_t *__restrict__ perm,
                                                    const uint64_t *__restrict__ sl_off,
                                                    const uint32_t *__restrict__ sl_len,
                                                    const T *__restrict__ eri,
                                                    const T *__restrict__ x, T *__restrict__ y, int odd) {
  typedef typename Pack<T, EB>::V V;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  V *W = reinterpret_cast<V *>(lds_raw);  // K ERI values per pair
  V *X = W + K;                           // nb ket values per pair, then a zero
  // byte offsets of an entry's two operands, from offsets pre-shifted for
  // 16-byte elements
  constexpr uint32_t vb = (uint32_t)sizeof(V);
  auto xoff = [](uint32_t w) { return vb >= 16 ? (w & 0x7fff0u) * (vb / 16) : (w & 0x7fff0u) / (16 / vb); };
  auto woff = [](uint32_t w) {
    return vb >= 16 ? ((w >> 15) & 0xfff0u) * (vb / 16) : ((w >> 15) & 0xfff0u) / (16 / vb);
  };
  // the kernel has no static LDS, so the dynamic array starts at LDS
  // address 0 (checked by the launcher) and an offset is an address
  typedef __attribute__((address_space(3))) const V lds_v;
  auto at = [](uint32_t off) { return *reinterpret_cast<lds_v *>((size_t)off); };
  // sum_q w_q x_q
  auto dot = [](const V &w, const V &xv) {
    T t = lane_c<T, EB>(w, 0) * lane_c<T, EB>(xv, 0);
#pragma unroll
    for (int q = 1; q < EB; q++) t = fma(lane_c<T, EB>(w, q), lane_c<T, EB>(xv, q), t);
    return t;
  };
  // one entry's contribution from its staged values: a += s * sum_q W_q X_q
  // (one FMA with the signed ket value when EB == 1)
  auto entry_vals = [&](uint32_t w, const V &wv, const V &xv, T &a) {
    if (EB == 1)
      a = fma(lane_c<T, EB>(wv, 0), flip_sign(lane_c<T, EB>(xv, 0), w & 0x80000000u), a);
    else
      a += flip_sign(dot(wv, xv), w & 0x80000000u);
  };
  const int L1 = L + 1;
  const int pos = blockIdx.x;
  const int I = SYM ? (int)perm[pos] : blockIdx.x;
  const int base = (SYM ? (pos & ~63) : 0) + blockIdx.y * (1024 * NT);
  if (SYM && base >= nslots) return;  // uniform per block, before any barrier
  const int tid = base + threadIdx.x;
  T acc[NT];
  uint64_t occ[NT];
  int rr[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) {
    acc[t] = T(0);
    const int slot = tid + t * 1024;
    const uint32_t r = slot < nslots ? perm[slot] : 0xffffffffu;
    rr[t] = r != 0xffffffffu ? (int)r : 0;
    occ[t] = r != 0xffffffffu ? b_str[r] : 0ull;  // padding slots: no dressing
  }
  const uint64_t occ_a = a_str[I];
  const uint64_t e0 = a_off[I], e1 = a_off[I + 1] + 1;  // + the spectator pair
  for (uint64_t e = e0; e < e1; e += EB) {
    // pair q of this round: kind 1 = alpha single, 2 = spectator, 0 = empty
    int kind[EB];
    const T *xsrc[EB];
    const T *vsrc[EB];
    T sgn[EB];
#pragma unroll
    for (int q = 0; q < EB; q++) {
      kind[q] = 0;
      xsrc[q] = x;
      vsrc[q] = eri;
      sgn[q] = T(1);
      if (e + q + 1 < e1) {
        const uint64_t ent = a_ent[e + q];
        kind[q] = 1;
        xsrc[q] = x + (size_t)ent_ket(ent) * nb;
        vsrc[q] = eri + (size_t)ent_pq(ent) * K;
        sgn[q] = ent_neg(ent) ? T(-1) : T(1);
      } else if (e + q + 1 == e1) {
        kind[q] = 2;
        xsrc[q] = x + (size_t)I * nb;
      }
    }
    // the ket rows are staged CH loop trips at a time (16 registers), all
    // loads of a chunk before its stores, so a chunk waits for one memory
    // latency rather than one per trip; the first chunk loads before the
    // barrier that ends the previous round's reads
    constexpr int CH = sizeof(V) >= 16 ? 4 : sizeof(V) == 8 ? 8 : 16;
    V xs[CH];
    auto load_x = [&](int i0) {
#pragma unroll
      for (int k = 0; k < CH; k++) {
        const int i = i0 + k * 1024;
#pragma unroll
        for (int q = 0; q < EB; q++) lane_of<T, EB>(xs[k], q) = (kind[q] && i < nb) ? xsrc[q][i] : T(0);
      }
    };
    auto store_x = [&](int i0) {
#pragma unroll
      for (int k = 0; k < CH; k++) {
        const int i = i0 + k * 1024;
        if (i <= nb) X[i] = xs[k];
      }
    };
    load_x(threadIdx.x);
    __syncthreads();
    store_x(threadIdx.x);
    for (int i0 = threadIdx.x + CH * 1024; i0 <= nb; i0 += CH * 1024) {
      load_x(i0);
      store_x(i0);
    }
    for (int i0 = threadIdx.x; i0 < K; i0 += 4 * 1024) {
      V ws[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int i = i0 + k * 1024;
#pragma unroll
        for (int q = 0; q < EB; q++) {
          T w = T(0);
          if (i < K && kind[q] == 1) w = sgn[q] * vsrc[q][i];
          else if (i < K && kind[q] == 2) w = spectator_val(eri, occ_a, L1, K, i);
          lane_of<T, EB>(ws[k], q) = w;
        }
      }
#pragma unroll
      for (int k = 0; k < 4; k++)
        if (i0 + k * 1024 < K) W[i0 + k * 1024] = ws[k];
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const int slot = tid + t * 1024;
      if (slot < nslots) {
        // a wavefront's 64 slots lie in one slice (slot bases are multiples
        // of 64), so the slice's group count and offset are wave-uniform
        const int sl = __builtin_amdgcn_readfirstlane(slot >> 6);
        const uint32_t ng = __builtin_amdgcn_readfirstlane(sl_len[sl]) / 4;
        // group g of this lane (4 entries, one 16-byte load) at p[64 g]
        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
        const u32x4 *p = reinterpret_cast<const u32x4 *>(ell + sl_off[sl]) + (slot & 63);
        T a = T(0);
        if (ng > 0) {
          // table words are loaded two groups ahead, so their L2 latency
          // overlaps two groups of LDS reads and FMAs; the loads run up to two
          // groups past the slice's end (into the next slice or the table's
          // two groups of padding) and those words are never used
          u32x4 g1 = p[0], g2 = p[64];
          typedef __attribute__((address_space(1))) const u32x4 gl_u4;
          gl_u4 *q = (gl_u4 *)(p + 128);  // group g + 2
          for (uint32_t g = 0; g < ng; g++) {
            const u32x4 cur = g1;
            g1 = g2;
            // an opaque address (an empty asm statement, no instruction):
            // otherwise the compiler re-loads each group where it is used
            // and the loads no longer run ahead
            asm volatile("" : "+v"(q));
            g2 = *q;
            q += 64;
            const uint32_t w4[4] = {cur.x, cur.y, cur.z, cur.w};
            // all eight LDS reads of the group issue before the first use
            V wv[4], xv[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
              wv[u] = at(woff(w4[u]));
              xv[u] = at(xoff(w4[u]));
            }
#pragma unroll
            for (int u = 0; u < 4; u++) entry_vals(w4[u], wv[u], xv[u], a);
          }
        }
        a += dot(X[rr[t]], diag_sum(W, occ[t], L1));
        acc[t] += a;
      }
    }
  }
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const int slot = tid + t * 1024;
    if (slot < nslots) {
      const T v = acc[t];
      const uint32_t r = perm[slot];
      if (SYM) {
        if (r != 0xffffffffu && slot > pos) {
          y[(size_t)I * nb + r] += v;
          y[(size_t)r * nb + I] += odd ? -v : v;
        } else if (r != 0xffffffffu && slot == pos && !odd) {
          y[(size_t)I * nb + r] += v;  // odd sector: the diagonal stays 0
        }
      } else {
        if (r != 0xffffffffu) y[(size_t)I * nb + r] += v;
      }
    }
  }
}

// Byte offset of the ket rows in k_mixed_sw's LDS image: the signed ERI
// triangle (2*Kt 16-byte elements, Kt <= 1024) lies below it, and the
// constant fits a DS instruction's 16-bit offset field.
constexpr uint32_t kSwKetBase = 32768;

// sum_{k in occ} v[tri(k,k)], tri(k,k) = k(k+3)/2 (the staged triangle's
// diagonal: the dressing sum of k_mixed_sw)
template <typename T, int EB>
__device__ __forceinline__ typename Pack<T, EB>::V diag_sum_tri(const typename Pack<T, EB>::V *v, uint64_t occ) {
  typename Pack<T, EB>::V s;
#pragma unroll
  for (int q = 0; q < EB; q++) lane_of<T, EB>(s, q) = T(0);
  while (occ) {
    const int k = __builtin_ctzll(occ);
    s += v[k * (k + 3) / 2];
    occ &= occ - 1;
  }
  return s;
}

// Mixed alpha-beta term with signed-triangle staging (SpinTables::ell3) and
// 16-byte pairs (EB = 4 FP32 or 2 FP64 values per LDS element).  Rounds,
// slots, slices, dressing and the SYM triangle are those of k_mixed_ell; the
// staged ERI row of each pair is kept over the triangle r >= s only (the ERI
// matrix being symmetric in its second index pair) and twice, with both
// signs: element t holds v[tri_rs[t]] and element Kt + t holds -v[tri_rs[t]].
// A beta single's sign is then part of its ERI address, and one entry costs
// one mask, one shift, two 16-byte LDS reads and EB fused multiply-adds into
// one of four running sums (no sign flip, no separate dot product).  The ket
// rows start at byte kSwKetBase, folded into the ket read's offset field.
template <typename T, int NT, bool SYM>
__global__ void __launch_bounds__(1024) k_mixed_sw(int nb, int K, int L, int Kt, int nslots,
                                                   const uint64_t *__restrict__ a_off,
                                                   const uint64_t *__restrict__ a_ent,
                                                   const uint64_t *__restrict__ a_str,
                                                   const uint64_t *__restrict__ b_str,
                                                   const uint32_t *__restrict__ ell,
                                                   const uint32_t *__restrict__ perm,
                                                   const uint64_t *__restrict__ sl_off,
                                                   const uint32_t *__restrict__ sl_len,
                                                   const uint16_t *__restrict__ tri_rs,
                                                   const T *__restrict__ eri,
                                                   const T *__restrict__ x, T *__restrict__ y, int odd) {
  constexpr int EB = 16 / sizeof(T);
  typedef typename Pack<T, EB>::V V;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  V *W = reinterpret_cast<V *>(lds_raw);               // 2*Kt signed ERI values per pair
  V *X = reinterpret_cast<V *>(lds_raw + kSwKetBase);  // nb ket values per pair, then a zero
  // the kernel has no static LDS, so the dynamic array starts at LDS
  // address 0 (checked by the launcher) and an offset is an address
  typedef __attribute__((address_space(3))) const V lds_v;
  auto at = [](uint32_t off) { return *reinterpret_cast<lds_v *>((size_t)off); };
  const int L1 = L + 1;
  const int pos = blockIdx.x;
  const int I = SYM ? (int)perm[pos] : blockIdx.x;
  const int base = (SYM ? (pos & ~63) : 0) + blockIdx.y * (1024 * NT);
  if (SYM && base >= nslots) return;  // uniform per block, before any barrier
  const int tid = base + threadIdx.x;
  T acc[NT];
  uint64_t occ[NT];
  int rr[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) {
    acc[t] = T(0);
    const int slot = tid + t * 1024;
    const uint32_t r = slot < nslots ? perm[slot] : 0xffffffffu;
    rr[t] = r != 0xffffffffu ? (int)r : 0;
    occ[t] = r != 0xffffffffu ? b_str[r] : 0ull;  // padding slots: no dressing
  }
  const uint64_t occ_a = a_str[I];
  const uint64_t e0 = a_off[I], e1 = a_off[I + 1] + 1;  // + the spectator pair
  for (uint64_t e = e0; e < e1; e += EB) {
    // pair q of this round: kind 1 = alpha single, 2 = spectator, 0 = empty
    int kind[EB];
    const T *xsrc[EB];
    const T *vsrc[EB];
    T sgn[EB];
#pragma unroll
    for (int q = 0; q < EB; q++) {
      kind[q] = 0;
      xsrc[q] = x;
      vsrc[q] = eri;
      sgn[q] = T(1);
      if (e + q + 1 < e1) {
        const uint64_t ent = a_ent[e + q];
        kind[q] = 1;
        xsrc[q] = x + (size_t)ent_ket(ent) * nb;
        vsrc[q] = eri + (size_t)ent_pq(ent) * K;
        sgn[q] = ent_neg(ent) ? T(-1) : T(1);
      } else if (e + q + 1 == e1) {
        kind[q] = 2;
        xsrc[q] = x + (size_t)I * nb;
      }
    }
    // every staged value of the round is loaded before any is stored
    // (nb + 1 <= 8192 ket values: at most 8 per thread; Kt <= 1024 ERI
    // values: at most one), so the round waits for one memory latency
    // rather than one per loop trip
    V xs[8], ws;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int i = threadIdx.x + k * 1024;
#pragma unroll
      for (int q = 0; q < EB; q++) lane_of<T, EB>(xs[k], q) = (kind[q] && i < nb) ? xsrc[q][i] : T(0);
    }
    const bool wl = (int)threadIdx.x < Kt;
    {
      const int rs = wl ? tri_rs[threadIdx.x] : 0;
#pragma unroll
      for (int q = 0; q < EB; q++) {
        T w = T(0);
        if (wl && kind[q] == 1) w = sgn[q] * vsrc[q][rs];
        else if (wl && kind[q] == 2) w = spectator_val(eri, occ_a, L1, K, rs);
        lane_of<T, EB>(ws, q) = w;
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int i = threadIdx.x + k * 1024;
      if (i <= nb) X[i] = xs[k];
    }
    if (wl) {
      W[threadIdx.x] = ws;
      W[Kt + threadIdx.x] = -ws;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const int slot = tid + t * 1024;
      if (slot < nslots) {
        // a wavefront's 64 slots lie in one slice (slot bases are multiples
        // of 64), so the slice's group count and offset are wave-uniform
        const int sl = __builtin_amdgcn_readfirstlane(slot >> 6);
        const uint32_t ng = __builtin_amdgcn_readfirstlane(sl_len[sl]) / 4;
        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
        const u32x4 *p = reinterpret_cast<const u32x4 *>(ell + sl_off[sl]) + (slot & 63);
        T s4[4] = {T(0), T(0), T(0), T(0)};
        // one group: 8 LDS reads issued together, then 4*EB FMAs into four
        // independent running sums
        auto group = [&](const u32x4 &cur) {
          const uint32_t w4[4] = {cur.x, cur.y, cur.z, cur.w};
          V wv[4], xv[4];
#pragma unroll
          for (int u = 0; u < 4; u++) {
            wv[u] = at(w4[u] >> 17);
            xv[u] = at((w4[u] & 0x1fff0u) + kSwKetBase);
          }
#pragma unroll
          for (int u = 0; u < 4; u++)
#pragma unroll
            for (int qq = 0; qq < EB; qq++)
              s4[u] = fma(lane_c<T, EB>(wv[u], qq), lane_c<T, EB>(xv[u], qq), s4[u]);
        };
        if (ng > 0) {
          // two table-word buffers in turn: the words of group g + 2 load
          // into the buffer group g has just used, so each load has a whole
          // group of LDS reads and FMAs to land and no register copies are
          // needed.  The loads run at most one group past the slice (into
          // the next slice or the table's two groups of padding); those
          // words are never used.
          u32x4 b0 = p[0], b1 = p[64];
          typedef __attribute__((address_space(1))) const u32x4 gl_u4;
          gl_u4 *q = (gl_u4 *)(p + 128);
          uint32_t g = 0;
          for (; g + 2 <= ng; g += 2) {
            // an opaque address (an empty asm statement, no instruction):
            // otherwise the compiler re-loads each group where it is used
            asm volatile("" : "+v"(q));
            group(b0);
            b0 = q[0];
            group(b1);
            b1 = q[64];
            q += 128;
          }
          if (g < ng) group(b0);
        }
        const V dv = diag_sum_tri<T, EB>(W, occ[t]) * X[rr[t]];
        T a = (s4[0] + s4[1]) + (s4[2] + s4[3]);
#pragma unroll
        for (int qq = 0; qq < EB; qq++) a += lane_c<T, EB>(dv, qq);
        acc[t] += a;
      }
    }
  }
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const int slot = tid + t * 1024;
    if (slot < nslots) {
      const T v = acc[t];
      const uint32_t r = perm[slot];
      if (SYM) {
        if (r != 0xffffffffu && slot > pos) {
          y[(size_t)I * nb + r] += v;
          y[(size_t)r * nb + I] += odd ? -v : v;
        } else if (r != 0xffffffffu && slot == pos && !odd) {
          y[(size_t)I * nb + r] += v;  // odd sector: the diagonal stays 0
        }
      } else {
        if (r != 0xffffffffu) y[(size_t)I * nb + r] += v;
      }
    }
  }
}

// Mixed term for KV vectors at once (x_v = x + v*N): one ERI row and the KV
// ket rows x_v[J,:] staged per alpha single (then the spectator pair), so
// each sliced-ELL entry costs one table load and 1 + KV LDS reads for KV
// multiply-adds.  Dressing as in k_mixed_ell.
template <typename T, int NT, int KV>
__global__ void __launch_bounds__(1024) k_mixed_multi(int nb, int K, int L, int nslots, size_t N,
                                                      const uint64_t *__restrict__ a_off,
                                                      const uint64_t *__restrict__ a_ent,
                                                      const uint64_t *__restrict__ a_str,
                                                      const uint64_t *__restrict__ b_str,
                                                      const uint32_t *__restrict__ ell,
                                                      const uint32_t *__restrict__ perm,
                                                      const uint64_t *__restrict__ sl_off,
                                                      const uint32_t *__restrict__ sl_len,
                                                      const T *__restrict__ eri,
                                                      const T *__restrict__ x, T *__restrict__ y) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *vrow = reinterpret_cast<T *>(lds_raw);  // K + 1
  T *crow = vrow + K + 1;                    // KV rows of nb
  const int L1 = L + 1;
  const int I = blockIdx.x;
  const int tid = blockIdx.y * (1024 * NT) + threadIdx.x;
  T acc[KV][NT];
  uint64_t occ[NT];
  int rr[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) {
#pragma unroll
    for (int v = 0; v < KV; v++) acc[v][t] = T(0);
    const int slot = tid + t * 1024;
    const uint32_t r = slot < nslots ? perm[slot] : 0xffffffffu;
    rr[t] = r != 0xffffffffu ? (int)r : 0;
    occ[t] = r != 0xffffffffu ? b_str[r] : 0ull;
  }
  const uint64_t occ_a = a_str[I];
  const uint64_t e1 = a_off[I + 1];
  for (uint64_t e = a_off[I]; e <= e1; e++) {
    size_t J = I;
    __syncthreads();
    if (e < e1) {
      const uint64_t ent = a_ent[e];
      J = ent_ket(ent);
      const T sa = ent_neg(ent) ? T(-1) : T(1);
      const T *vsrc = eri + (size_t)ent_pq(ent) * K;
      for (int i = threadIdx.x; i < K; i += 1024) vrow[i] = sa * vsrc[i];
    } else {
      for (int i = threadIdx.x; i < K; i += 1024) vrow[i] = spectator_val(eri, occ_a, L1, K, i);
    }
    if (threadIdx.x == 0) vrow[K] = T(0);
#pragma unroll
    for (int v = 0; v < KV; v++)
      for (int i = threadIdx.x; i < nb; i += 1024) crow[v * nb + i] = x[v * N + J * nb + i];
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const int slot = tid + t * 1024;
      if (slot < nslots) {
        const int sl = slot >> 6;
        const uint32_t len = sl_len[sl];
        const uint32_t *p = ell + sl_off[sl] + (slot & 63);
        T a[KV];
#pragma unroll
        for (int v = 0; v < KV; v++) a[v] = T(0);
#pragma unroll 4
        for (uint32_t k = 0; k < len; k++) {
          const uint32_t w = p[64 * (size_t)k];
          const int jb = (int)(w & 0xffffu);
          const T vr = vrow[(w >> 16) & 0x7fffu];
          const T sv = (w >> 31) ? -vr : vr;
#pragma unroll
          for (int v = 0; v < KV; v++) a[v] += sv * crow[v * nb + jb];
        }
        const T d = diag_sum(vrow, occ[t], L1);
#pragma unroll
        for (int v = 0; v < KV; v++) acc[v][t] += a[v] + crow[v * nb + rr[t]] * d;
      }
    }
  }
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const int slot = tid + t * 1024;
    if (slot < nslots) {
      const uint32_t r = perm[slot];
      if (r != 0xffffffffu) {
#pragma unroll
        for (int v = 0; v < KV; v++) y[v * N + (size_t)I * nb + r] += acc[v][t];
      }
    }
  }
}

// Mixed term when one ket row does not fit in LDS: the row is staged in
// nparts pieces of h kets and each piece is walked with its own sliced-ELL
// table (SpinTables::pell, kets relative to the piece); the sign-folded ERI
// row (or the spectator row, last) stays staged for all pieces, and the
// dressing of beta row r is added with the piece that holds ket r.
struct PartTables {
  const uint32_t *ell[8];
  const uint64_t *off[8];
  const uint32_t *len[8];
};

template <typename T, int NT, bool SYM = false>
__global__ void __launch_bounds__(1024) k_mixed_parts(int nb, int K, int L, int nslots, int nparts,
                                                      int h, const uint64_t *__restrict__ a_off,
                                                      const uint64_t *__restrict__ a_ent,
                                                      const uint64_t *__restrict__ a_str,
                                                      const uint64_t *__restrict__ b_str,
                                                      PartTables pt, const uint32_t *__restrict__ perm,
                                                      const T *__restrict__ eri,
                                                      const T *__restrict__ x, T *__restrict__ y,
                                                      int odd) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *crow = reinterpret_cast<T *>(lds_raw);
  T *vrow = crow + h;
  const int L1 = L + 1;
  const int pos = blockIdx.x;
  const int I = SYM ? (int)perm[pos] : blockIdx.x;
  // SYM: slots from pos's slice on (as in k_mixed_ell); the row's slots are
  // split over gridDim.y blocks of 1024*NT slots each
  const int base = (SYM ? (pos & ~63) : 0) + blockIdx.y * (1024 * NT);
  if (base >= nslots) return;  // uniform per block, before any barrier
  const int tid = base + threadIdx.x;
  T acc[NT];
  uint64_t occ[NT];
  int rr[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) {
    acc[t] = T(0);
    const int slot = tid + t * 1024;
    const uint32_t r = slot < nslots ? perm[slot] : 0xffffffffu;
    rr[t] = r != 0xffffffffu ? (int)r : 0;
    occ[t] = r != 0xffffffffu ? b_str[r] : 0ull;
  }
  const uint64_t occ_a = a_str[I];
  const uint64_t e1 = a_off[I + 1];
  for (uint64_t e = a_off[I]; e <= e1; e++) {
    const bool spect = e == e1;
    size_t J = I;
    __syncthreads();
    // staged rows go CH loop trips at a time (16 registers), all loads of a
    // chunk before its stores: one memory latency per chunk, not per trip
    constexpr int CH = sizeof(T) == 8 ? 8 : 16;
    T st[CH];
    if (!spect) {
      const uint64_t ent = a_ent[e];
      J = ent_ket(ent);
      const T sa = ent_neg(ent) ? T(-1) : T(1);
      const T *vsrc = eri + (size_t)ent_pq(ent) * K;
      for (int i0 = threadIdx.x; i0 < K; i0 += CH * 1024) {
#pragma unroll
        for (int k = 0; k < CH; k++) st[k] = i0 + k * 1024 < K ? sa * vsrc[i0 + k * 1024] : T(0);
#pragma unroll
        for (int k = 0; k < CH; k++)
          if (i0 + k * 1024 < K) vrow[i0 + k * 1024] = st[k];
      }
    } else {
      for (int i = threadIdx.x; i < K; i += 1024) vrow[i] = spectator_val(eri, occ_a, L1, K, i);
    }
    if (threadIdx.x == 0) vrow[K] = T(0);
    for (int p = 0; p < nparts; p++) {
      const int lo = p * h;
      const int len_p = (nb - lo) < h ? (nb - lo) : h;
      const T *src = x + J * nb + lo;
      // the part's first chunk loads before the barrier that ends the
      // previous part's reads
#pragma unroll
      for (int k = 0; k < CH; k++) {
        const int i = threadIdx.x + k * 1024;
        st[k] = i < len_p ? src[i] : T(0);
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < CH; k++) {
        const int i = threadIdx.x + k * 1024;
        if (i < len_p) crow[i] = st[k];
      }
      for (int i0 = threadIdx.x + CH * 1024; i0 < len_p; i0 += CH * 1024) {
#pragma unroll
        for (int k = 0; k < CH; k++) st[k] = i0 + k * 1024 < len_p ? src[i0 + k * 1024] : T(0);
#pragma unroll
        for (int k = 0; k < CH; k++)
          if (i0 + k * 1024 < len_p) crow[i0 + k * 1024] = st[k];
      }
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
          if ((unsigned)(rr[t] - lo) < (unsigned)len_p)
            a += crow[rr[t] - lo] * diag_sum(vrow, occ[t], L1);
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
      if (SYM) {
        if (r != 0xffffffffu && slot > pos) {
          y[(size_t)I * nb + r] += acc[t];
          y[(size_t)r * nb + I] += odd ? -acc[t] : acc[t];
        } else if (r != 0xffffffffu && slot == pos && !odd) {
          y[(size_t)I * nb + r] += acc[t];
        }
      } else {
        if (r != 0xffffffffu) y[(size_t)I * nb + r] += acc[t];
      }
    }
  }
}

// Mixed term with the beta singles in CSR form (used when the packed
// sliced-ELL table cannot hold the indices): x[J,:] and V row in LDS, then
// the spectator pair; dressing as in k_mixed_ell.
template <typename T, int NT>
__global__ void __launch_bounds__(1024) k_mixed_lds(int nb, int K, int L,
                                                    const uint64_t *__restrict__ a_off,
                                                    const uint64_t *__restrict__ a_ent,
                                                    const uint64_t *__restrict__ a_str,
                                                    const uint64_t *__restrict__ b_str,
                                                    const uint64_t *__restrict__ b_off,
                                                    const uint64_t *__restrict__ b_ent,
                                                    const T *__restrict__ eri,
                                                    const T *__restrict__ x, T *__restrict__ y) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *crow = reinterpret_cast<T *>(lds_raw);
  T *vrow = crow + nb;
  const int L1 = L + 1;
  const int I = blockIdx.x;
  const int tid = threadIdx.x;
  T acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) acc[t] = T(0);
  const uint64_t occ_a = a_str[I];
  const uint64_t e1 = a_off[I + 1];
  for (uint64_t e = a_off[I]; e <= e1; e++) {
    size_t J = I;
    __syncthreads();
    if (e < e1) {
      const uint64_t ent = a_ent[e];
      J = ent_ket(ent);
      const T sa = ent_neg(ent) ? T(-1) : T(1);
      const T *vsrc = eri + (size_t)ent_pq(ent) * K;
      for (int i = tid; i < K; i += 1024) vrow[i] = sa * vsrc[i];
    } else {
      for (int i = tid; i < K; i += 1024) vrow[i] = spectator_val(eri, occ_a, L1, K, i);
    }
    for (int i = tid; i < nb; i += 1024) crow[i] = x[J * nb + i];
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const int ib = tid + t * 1024;
      if (ib < nb) {
        T a = crow[ib] * diag_sum(vrow, b_str[ib], L1);
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
__global__ void __launch_bounds__(256) k_mixed_global(int nb, int nchunk, int K, int L,
                                                      const uint64_t *__restrict__ a_off,
                                                      const uint64_t *__restrict__ a_ent,
                                                      const uint64_t *__restrict__ a_str,
                                                      const uint64_t *__restrict__ b_str,
                                                      const uint64_t *__restrict__ b_off,
                                                      const uint64_t *__restrict__ b_ent,
                                                      const T *__restrict__ eri,
                                                      const T *__restrict__ x, T *__restrict__ y) {
  const int I = blockIdx.x / nchunk;
  const int ib = (blockIdx.x % nchunk) * blockDim.x + threadIdx.x;
  if (ib >= nb) return;
  const int L1 = L + 1;
  const uint64_t f0 = b_off[ib], f1 = b_off[ib + 1];
  const uint64_t occ_b = b_str[ib];
  T acc = T(0);
  for (uint64_t e = a_off[I]; e < a_off[I + 1]; e++) {
    const uint64_t ent = a_ent[e];
    const T *xr = x + (size_t)ent_ket(ent) * nb;
    const T *vr = eri + (size_t)ent_pq(ent) * K;
    T a = xr[ib] * diag_sum(vr, occ_b, L1);
    for (uint64_t f = f0; f < f1; f++) {
      const uint64_t be = b_ent[f];
      const T v = vr[ent_pq(be)] * xr[ent_ket(be)];
      a += ent_neg(be) ? -v : v;
    }
    acc += ent_neg(ent) ? -a : a;
  }
  // spectator: beta singles dressed by the alpha string of I
  const uint64_t occ_a = a_str[I];
  const T *xr = x + (size_t)I * nb;
  for (uint64_t f = f0; f < f1; f++) {
    const uint64_t be = b_ent[f];
    const T v = spectator_val(eri, occ_a, L1, K, (int)ent_pq(be)) * xr[ent_ket(be)];
    acc += ent_neg(be) ? -v : v;
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

// Dense same-spin block from the tables: H[r*n + ket] = value (zeroed first).
template <typename T>
__global__ void __launch_bounds__(256) k_densify(int n, const uint64_t *__restrict__ s_off,
                                                 const uint64_t *__restrict__ s_ent,
                                                 const double *__restrict__ s_val,
                                                 const uint64_t *__restrict__ d_off,
                                                 const uint32_t *__restrict__ d_ket,
                                                 const double *__restrict__ d_val, T *__restrict__ H) {
  for (int r = blockIdx.x; r < n; r += gridDim.x) {
    T *row = H + (size_t)r * n;
    for (uint64_t e = s_off[r] + threadIdx.x; e < s_off[r + 1]; e += blockDim.x)
      row[ent_ket(s_ent[e])] = (T)s_val[e];
    for (uint64_t e = d_off[r] + threadIdx.x; e < d_off[r + 1]; e += blockDim.x)
      row[d_ket[e]] = (T)d_val[e];
  }
}

// Spin-flip symmetric form of the same-spin terms for equal string sets:
// y = hii*x + (G + G^T) with G = H X (n x n row-major), 64x64 tiles, G^T
// through LDS; for symmetric x the result is bitwise symmetric.  For the
// odd sector (x = -x^T, odd != 0) X H = -G^T, so y = hii*x + (G - G^T),
// bitwise antisymmetric with a zero diagonal.
// gscale (device, may be null): G holds 2^(ex+eh) times the product (the
// split-FP16 GEMM), and gscale[0] * gscale[1] = 2^-(ex+eh) undoes it exactly.
template <typename T>
__global__ void __launch_bounds__(256) k_sym_combine(int n, const T *__restrict__ hii, const T *__restrict__ x,
                                                     const T *__restrict__ G, T *__restrict__ y,
                                                     const float *__restrict__ gscale, int odd) {
  const T s = gscale ? (T)(gscale[0] * gscale[1]) : T(1);
  __shared__ T tile[64][65];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int c0 = blockIdx.x * 64, r0 = blockIdx.y * 64;
  for (int k = ty; k < 64; k += 4) {
    const int gr = c0 + k, gc = r0 + tx;
    if (gr < n && gc < n) tile[k][tx] = G[(size_t)gr * n + gc];
  }
  __syncthreads();
  for (int k = ty; k < 64; k += 4) {
    const int r = r0 + k, cc = c0 + tx;
    if (r < n && cc < n) {
      const size_t o = (size_t)r * n + cc;
      // G + G^T commutes (bitwise symmetric); G - G^T is bitwise antisymmetric
      y[o] = hii[o] * x[o] + (odd ? G[o] - tile[tx][k] : G[o] + tile[tx][k]) * s;
    }
  }
}

// Split-FP16 operands of the symmetric FP32 same-spin GEMM.  An FP32 array
// v is written as 2^-e (hi + lo) with hi = fp16(v 2^e), lo = fp16(v 2^e - hi)
// and 2^e putting max|v| 2^e in [2^14, 2^15): about 22 significant bits
// down to an absolute floor of 2^-25 in scaled units.  Products of FP16
// values are exact in FP32 and the FP16 MFMA GEMMs accumulate in FP32, so
// X H = 2^-(ex+eh) (Xh Hh + Xh Hl + Xl Hh) up to the dropped Xl Hl
// (~2^-22 relative) and FP32 summation.
constexpr int kSplitParts = 1024;

__device__ __forceinline__ float block_max256(float m) {
  __shared__ float wmax[4];
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
  __syncthreads();
  return fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
}

// part[b] = max |v[i]| over block b's grid-stride share (gridDim.x blocks).
__global__ void __launch_bounds__(256) k_absmax_part(size_t n, const float *__restrict__ v,
                                                     float *__restrict__ part) {
  float m = 0.f;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    m = fmaxf(m, fabsf(v[i]));
  m = block_max256(m);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// hi/lo halves of v with the scale taken from the np block maxima; block 0
// stores 2^-e at *inv.  e is capped at 100 (vectors below 2^-85 in max
// norm keep the absolute floor 2^-125).  hi2 (may be null) receives a second
// copy of hi.
__global__ void __launch_bounds__(256) k_split_f16(size_t n, const float *__restrict__ v,
                                                   const float *__restrict__ part, int np,
                                                   _Float16 *__restrict__ hi, _Float16 *__restrict__ lo,
                                                   _Float16 *__restrict__ hi2, float *__restrict__ inv) {
  float m = 0.f;
  for (int i = threadIdx.x; i < np; i += 256) m = fmaxf(m, part[i]);
  m = block_max256(m);
  int k = 0;
  if (m > 0.f) (void)frexpf(m, &k);  // m = f 2^k, f in [0.5, 1)
  const int e = m > 0.f ? std::min(15 - k, 100) : 0;
  const float sc = ldexpf(1.f, e);
  if (blockIdx.x == 0 && threadIdx.x == 0) *inv = ldexpf(1.f, -e);
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const float t = v[i] * sc;
    const _Float16 h = (_Float16)t;
    hi[i] = h;
    lo[i] = (_Float16)(t - (float)h);
    if (hi2) hi2[i] = h;
  }
}

// v = (v + p v^T)/2 on an n x n row-major matrix, p = +1 (even sector) or
// -1 (odd sector: the diagonal becomes 0); block (bx, by), bx <= by, owns
// the tile pair (by, bx) / (bx, by).
__global__ void __launch_bounds__(256) k_symmetrize(int n, double *__restrict__ v, double p) {
  if (blockIdx.x > blockIdx.y) return;
  __shared__ double A[64][65], B[64][65];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int R0 = blockIdx.y * 64, C0 = blockIdx.x * 64;
  for (int k = ty; k < 64; k += 4) {
    A[k][tx] = (R0 + k < n && C0 + tx < n) ? v[(size_t)(R0 + k) * n + C0 + tx] : 0.0;
    B[k][tx] = (C0 + k < n && R0 + tx < n) ? v[(size_t)(C0 + k) * n + R0 + tx] : 0.0;
  }
  __syncthreads();
  for (int k = ty; k < 64; k += 4) {
    if (R0 + k < n && C0 + tx < n) v[(size_t)(R0 + k) * n + C0 + tx] = 0.5 * (A[k][tx] + p * B[tx][k]);
    if (C0 + k < n && R0 + tx < n) v[(size_t)(C0 + k) * n + R0 + tx] = 0.5 * (B[k][tx] + p * A[tx][k]);
  }
}

// part[2*b] = max |v - p v^T| and part[2*b+1] = max |v| over tile pair b.
__global__ void __launch_bounds__(256) k_sym_defect(int n, const double *__restrict__ v, double p,
                                                    double *__restrict__ part) {
  __shared__ double A[64][65];
  __shared__ double red[2][256];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int R0 = blockIdx.y * 64, C0 = blockIdx.x * 64;
  for (int k = ty; k < 64; k += 4)
    A[k][tx] = (C0 + k < n && R0 + tx < n) ? v[(size_t)(C0 + k) * n + R0 + tx] : 0.0;
  __syncthreads();
  double d = 0.0, m = 0.0;
  for (int k = ty; k < 64; k += 4)
    if (R0 + k < n && C0 + tx < n) {
      const double a = v[(size_t)(R0 + k) * n + C0 + tx];
      d = fmax(d, fabs(a - p * A[tx][k]));
      m = fmax(m, fabs(a));
    }
  red[0][threadIdx.x] = d;
  red[1][threadIdx.x] = m;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      red[0][threadIdx.x] = fmax(red[0][threadIdx.x], red[0][threadIdx.x + st]);
      red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + st]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const size_t b = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
    part[2 * b] = red[0][0];
    part[2 * b + 1] = red[1][0];
  }
}

// y = hii * x (the diagonal term of the GEMM form).
template <typename T>
__global__ void __launch_bounds__(256) k_diag_mul(size_t n, const T *__restrict__ hii,
                                                  const T *__restrict__ x, T *__restrict__ y) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    y[i] = hii[i] * x[i];
}

// ------------------------------------------------------------------ host side
template <typename T>
struct SigmaOperands {
  const T *s_val_a, *d_val_a, *s_val_b, *d_val_b, *eri, *hii;
  T *T1, *T2;
};

static SigmaOperands<double> operands(Context &c, double *) {
  return {c.A.s_val, c.A.d_val, c.B.s_val, c.B.d_val, c.eri, c.hii, c.T1, c.T2};
}
static SigmaOperands<float> operands(Context &c, float *) {
  return {c.f32.s_val_a, c.f32.d_val_a, c.f32.s_val_b, c.f32.d_val_b,
          c.f32.eri,     c.f32.hii,     c.f32.T1,      c.f32.T2};
}

template <typename T>
static void sigma_same(Context &c, const SpinTables &t, const T *s_val, const T *d_val, int nrows,
                       int M, const T *hii, const T *x, T *y, bool diag) {
  const int nch = (int)cdiv(M, 64), nrb = (int)cdiv(nrows, 4);
  const size_t nblk = (size_t)8 * cdiv(nch, 8) * nrb;
  if (nblk >= (1ull << 31)) throw Error(3, "sigma: grid too large for one launch");
  if (diag)
    hipLaunchKernelGGL((k_same_xcd<T, true>), dim3((unsigned)nblk), dim3(256), 0, c.stream, nrows, M,
                       nch, nrb, t.s_off, t.s_ent, s_val, t.d_off, t.d_ket, d_val, hii, x, y);
  else
    hipLaunchKernelGGL((k_same_xcd<T, false>), dim3((unsigned)nblk), dim3(256), 0, c.stream, nrows, M,
                       nch, nrb, t.s_off, t.s_ent, s_val, t.d_off, t.d_ket, d_val, hii, x, y);
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

// Rows of the mixed term are split over ceil(slots / (1024*NT)) blocks.
template <typename T, int NT, int EB, bool SYM>
static void launch_mixed_ell(Context &c, const T *eri, const T *x, T *y, size_t lds) {
  auto fn = k_mixed_ell<T, NT, EB, SYM>;
  static bool lds_base_checked = false;
  if (!lds_base_checked) {
    // k_mixed_ell addresses LDS by byte offset from 0: no static LDS allowed
    hipFuncAttributes attr;
    SBDX_HIP(hipFuncGetAttributes(&attr, (const void *)fn));
    if (attr.sharedSizeBytes != 0) throw Error(2, "k_mixed_ell: static LDS would offset the dynamic array");
    lds_base_checked = true;
  }
  SBDX_HIP(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const int nslots = c.B.nslice * 64;
  const unsigned nsplit = cdiv(nslots, 1024 * NT);
  hipLaunchKernelGGL(fn, dim3((unsigned)c.na, nsplit), dim3(1024), lds, c.stream, (int)c.nb, c.K, c.L,
                     nslots, c.A.s_off, c.A.s_ent, c.A.str, c.B.str, c.B.ell2, c.B.perm, c.B.sl_off2,
                     c.B.sl_len2, eri, x, y, (SYM && c.sym_par < 0) ? 1 : 0);
}

// EB pairs staged per round, NT slots per thread (1, 2, 4 or 8); false when
// the staging does not fit in LDS.
template <typename T, int EB, bool SYM = false>
static bool mixed_ell_dispatch(Context &c, int nt, const T *eri, const T *x, T *y) {
  const size_t lds = sizeof(T) * EB * ((size_t)c.nb + c.K + 1);
  if (lds > c.lds_limit || !c.B.has_ell2) return false;
  if (nt <= 1) launch_mixed_ell<T, 1, EB, SYM>(c, eri, x, y, lds);
  else if (nt <= 2) launch_mixed_ell<T, 2, EB, SYM>(c, eri, x, y, lds);
  else if (nt <= 4) launch_mixed_ell<T, 4, EB, SYM>(c, eri, x, y, lds);
  else launch_mixed_ell<T, 8, EB, SYM>(c, eri, x, y, lds);
  return true;
}

// Staging width code of the signed-triangle kernel in the autotuner's
// (width, slots) choice: 16-byte pairs, 4 FP32 or 2 FP64 values.
constexpr int kEbSigned = 16;

template <typename T, int NT, bool SYM>
static void launch_mixed_sw(Context &c, const T *eri, const T *x, T *y, size_t lds) {
  auto fn = k_mixed_sw<T, NT, SYM>;
  static bool lds_base_checked = false;
  if (!lds_base_checked) {
    hipFuncAttributes attr;
    SBDX_HIP(hipFuncGetAttributes(&attr, (const void *)fn));
    if (attr.sharedSizeBytes != 0) throw Error(2, "k_mixed_sw: static LDS would offset the dynamic array");
    lds_base_checked = true;
  }
  SBDX_HIP(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const int nslots = c.B.nslice * 64;
  const unsigned nsplit = cdiv(nslots, 1024 * NT);
  hipLaunchKernelGGL(fn, dim3((unsigned)c.na, nsplit), dim3(1024), lds, c.stream, (int)c.nb, c.K, c.L, c.Kt,
                     nslots, c.A.s_off, c.A.s_ent, c.A.str, c.B.str, c.B.ell3, c.B.perm, c.B.sl_off2,
                     c.B.sl_len2, c.tri_rs, eri, x, y, (SYM && c.sym_par < 0) ? 1 : 0);
}

// The signed-triangle kernel, NT slots per thread; false when its table was
// not built or its LDS image does not fit.
template <typename T, bool SYM = false>
static bool mixed_sw_dispatch(Context &c, int nt, const T *eri, const T *x, T *y) {
  const size_t lds = kSwKetBase + (size_t)16 * ((size_t)c.nb + 1);
  if (!c.B.has_ell3 || (size_t)32 * c.Kt > kSwKetBase || c.nb + 1 > 8192 || lds > c.lds_limit) return false;
  if (nt <= 1) launch_mixed_sw<T, 1, SYM>(c, eri, x, y, lds);
  else if (nt <= 2) launch_mixed_sw<T, 2, SYM>(c, eri, x, y, lds);
  else if (nt <= 4) launch_mixed_sw<T, 4, SYM>(c, eri, x, y, lds);
  else launch_mixed_sw<T, 8, SYM>(c, eri, x, y, lds);
  return true;
}

template <typename T, int NT>
static void launch_mixed_lds(Context &c, const T *eri, const T *x, T *y, size_t lds) {
  auto fn = k_mixed_lds<T, NT>;
  SBDX_HIP(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(fn, dim3((unsigned)c.na), dim3(1024), lds, c.stream, (int)c.nb, c.K, c.L,
                     c.A.s_off, c.A.s_ent, c.A.str, c.B.str, c.B.s_off, c.B.s_ent, eri, x, y);
}

template <typename T, int NT, bool SYM = false>
static void launch_mixed_parts(Context &c, const T *eri, const T *x, T *y, size_t lds) {
  auto fn = k_mixed_parts<T, NT, SYM>;
  SBDX_HIP(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  PartTables pt{};
  for (int p = 0; p < c.B.nparts; p++) {
    pt.ell[p] = c.B.pell[p];
    pt.off[p] = c.B.psl_off[p];
    pt.len[p] = c.B.psl_len[p];
  }
  const int nslots = c.B.nslice * 64;
  const unsigned nsplit = cdiv(nslots, 1024 * NT);
  hipLaunchKernelGGL(fn, dim3((unsigned)c.na, nsplit), dim3(1024), lds, c.stream, (int)c.nb, c.K, c.L, nslots,
                     c.B.nparts, c.B.part_h, c.A.s_off, c.A.s_ent, c.A.str, c.B.str, pt, c.B.perm, eri, x,
                     y, (SYM && c.sym_par < 0) ? 1 : 0);
}

template <typename T, bool SYM = false>
static void launch_mixed_parts_nt(Context &c, int nt, const T *eri, const T *x, T *y, size_t lds) {
  if (nt <= 1) launch_mixed_parts<T, 1, SYM>(c, eri, x, y, lds);
  else if (nt <= 2) launch_mixed_parts<T, 2, SYM>(c, eri, x, y, lds);
  else if (nt <= 4) launch_mixed_parts<T, 4, SYM>(c, eri, x, y, lds);
  else if (nt <= 8) launch_mixed_parts<T, 8, SYM>(c, eri, x, y, lds);
  else if (nt <= 16) launch_mixed_parts<T, 16, SYM>(c, eri, x, y, lds);
  else launch_mixed_parts<T, 32, SYM>(c, eri, x, y, lds);
}

static double *scratch_vec(Context &c, const double *) { return c.T2; }
static float *scratch_vec(Context &c, const float *) { return c.f32.T2; }

// Slots per thread NT of the row-parts kernel (a row's slots split over
// ceil(slots / (1024 NT)) blocks, each staging the same rows): every power
// of two up to the one that covers a row in one block (at most 32) runs
// twice on this x into a scratch vector on the first product of this
// precision and form, HIP events time the second run, and the first within
// 3% of the fastest is kept (large NT saves staging, small NT registers).
// SBDX_PARTS_NT=1|2|4|8|16|32 fixes it.
template <typename T, bool SYM = false>
static int parts_nt(Context &c, const T *eri, const T *x, size_t lds) {
  const int pi = (sizeof(T) == 4 ? 1 : 0) + (SYM ? 2 : 0);
  if (c.parts_nt[pi] > 0) return c.parts_nt[pi];
  const int nt_max = (int)cdiv((size_t)c.B.nslice * 64, 1024);
  if (const char *ev = std::getenv("SBDX_PARTS_NT")) {
    const int v = std::atoi(ev);
    if (v != 1 && v != 2 && v != 4 && v != 8 && v != 16 && v != 32)
      throw Error(1, "SBDX_PARTS_NT must be 1, 2, 4, 8, 16 or 32");
    return c.parts_nt[pi] = v;
  }
  T *scratch = scratch_vec(c, x);
  hipEvent_t e0, e1;
  SBDX_HIP(hipEventCreate(&e0));
  SBDX_HIP(hipEventCreate(&e1));
  std::vector<std::pair<int, float>> timed;
  for (int nt : {1, 2, 4, 8, 16, 32}) {
    if (nt > 1 && nt / 2 >= nt_max) break;  // a smaller NT already covers the row
    float ms = 0.f;
    for (int rep = 0; rep < 2; rep++) {
      SBDX_HIP(hipEventRecord(e0, c.stream));
      launch_mixed_parts_nt<T, SYM>(c, nt, eri, x, scratch, lds);
      SBDX_HIP(hipEventRecord(e1, c.stream));
      SBDX_HIP(hipEventSynchronize(e1));
      SBDX_HIP(hipEventElapsedTime(&ms, e0, e1));
    }
    check_launch("k_mixed_parts");
    timed.push_back({nt, ms});
  }
  SBDX_HIP(hipEventDestroy(e0));
  SBDX_HIP(hipEventDestroy(e1));
  float fastest = timed[0].second;
  for (const auto &t : timed) fastest = std::min(fastest, t.second);
  for (const auto &t : timed)
    if (t.second <= 1.03f * fastest) return c.parts_nt[pi] = t.first;
  return c.parts_nt[pi] = timed.back().first;
}

// Ket row staged in parts: smallest part count (<= 8) whose piece and ERI row fit.
template <typename T, bool SYM = false>
static bool mixed_parts_dispatch(Context &c, const T *eri, const T *x, T *y) {
  // Parts already built (e.g. for the other precision) are kept when they fit.
  const int first = c.B.nparts >= 2 ? c.B.nparts : 2;
  for (int P = first; P <= 8; P++) {
    const size_t h = (c.nb + P - 1) / P;
    const size_t lds = sizeof(T) * (h + c.K + 1);
    if (lds > c.lds_limit) continue;
    build_singles_ell_parts(c, c.B, P);
    launch_mixed_parts_nt<T, SYM>(c, parts_nt<T, SYM>(c, eri, x, lds), eri, x, y, lds);
    return true;
  }
  return false;
}

template <bool SYM = false>
static bool mixed_ell_cfg(Context &c, int eb, int nt, const double *eri, const double *x, double *y) {
  return eb == kEbSigned ? mixed_sw_dispatch<double, SYM>(c, nt, eri, x, y)
       : eb == 4 ? mixed_ell_dispatch<double, 4, SYM>(c, nt, eri, x, y)
       : eb == 2 ? mixed_ell_dispatch<double, 2, SYM>(c, nt, eri, x, y)
                 : mixed_ell_dispatch<double, 1, SYM>(c, nt, eri, x, y);
}
template <bool SYM = false>
static bool mixed_ell_cfg(Context &c, int eb, int nt, const float *eri, const float *x, float *y) {
  return eb == kEbSigned ? mixed_sw_dispatch<float, SYM>(c, nt, eri, x, y)
       : eb == 4 ? mixed_ell_dispatch<float, 4, SYM>(c, nt, eri, x, y)
       : eb == 2 ? mixed_ell_dispatch<float, 2, SYM>(c, nt, eri, x, y)
                 : mixed_ell_dispatch<float, 1, SYM>(c, nt, eri, x, y);
}
// Autotuner for the sliced-ELL mixed kernel: every (staging width EB in
// {1, 2, 4, kEbSigned (k_mixed_sw)}, slots per thread NT in {1, 2, 4, 8})
// whose LDS footprint fits
// and that does not cover fewer than half of a row's slots per thread
// budget runs twice on this x into a scratch vector (the kernel
// accumulates, so y is untouched); HIP events on the stream time the second
// run and the fastest is kept for this precision (-1: nothing fits).
// SBDX_MIXED_EB=1|2|4|16 and SBDX_MIXED_NT=1|2|4|8 fix either coordinate.
template <typename T, bool SYM = false>
static int mixed_cfg(Context &c, const T *eri, const T *x) {
  const int pi = (sizeof(T) == 4 ? 1 : 0) + (SYM ? 2 : 0);
  if (c.mixed_eb[pi] != 0) return c.mixed_eb[pi];
  const int nt_max = (int)cdiv((size_t)c.B.nslice * 64, 1024);
  const char *env_eb = std::getenv("SBDX_MIXED_EB");
  const char *env_nt = std::getenv("SBDX_MIXED_NT");
  const int fix_eb = env_eb ? std::atoi(env_eb) : 0, fix_nt = env_nt ? std::atoi(env_nt) : 0;
  if (env_eb && fix_eb != 1 && fix_eb != 2 && fix_eb != 4 && fix_eb != kEbSigned)
    throw Error(1, "SBDX_MIXED_EB must be 1, 2, 4 or 16");
  if (env_nt && fix_nt != 1 && fix_nt != 2 && fix_nt != 4 && fix_nt != 8)
    throw Error(1, "SBDX_MIXED_NT must be 1, 2, 4 or 8");
  T *scratch = scratch_vec(c, x);
  hipEvent_t e0, e1;
  SBDX_HIP(hipEventCreate(&e0));
  SBDX_HIP(hipEventCreate(&e1));
  // every timed configuration, in the fixed order (the signed-triangle
  // kernel, then eb descending; nt ascending)
  std::vector<std::array<float, 3>> timed;
  for (int eb : {kEbSigned, 4, 2, 1}) {
    if (fix_eb && eb != fix_eb) continue;
    for (int nt : {1, 2, 4, 8}) {
      if (fix_nt && nt != fix_nt) continue;
      if (!fix_nt && nt > 1 && (nt / 2) >= nt_max) continue;  // one block already covers the row
      bool ok = true;
      float ms = 0.f;
      for (int rep = 0; rep < 2 && ok; rep++) {
        SBDX_HIP(hipEventRecord(e0, c.stream));
        ok = mixed_ell_cfg<SYM>(c, eb, nt, eri, x, scratch);
        SBDX_HIP(hipEventRecord(e1, c.stream));
        SBDX_HIP(hipEventSynchronize(e1));
        SBDX_HIP(hipEventElapsedTime(&ms, e0, e1));
      }
      if (!ok) continue;
      check_launch("k_mixed_ell");
      timed.push_back({(float)eb, (float)nt, ms});
    }
  }
  // the first configuration within 3% of the fastest: timings that differ by
  // less than run-to-run noise then give the same choice (and the same
  // summation order) in every run
  int best_eb = -1, best_nt = 0;
  float best_ms = 0.f;
  if (!timed.empty()) {
    float fastest = timed[0][2];
    for (const auto &t : timed) fastest = std::min(fastest, t[2]);
    for (const auto &t : timed)
      if (t[2] <= 1.03f * fastest) {
        best_eb = (int)t[0];
        best_nt = (int)t[1];
        best_ms = t[2];
        break;
      }
  }
  SBDX_HIP(hipEventDestroy(e0));
  SBDX_HIP(hipEventDestroy(e1));
  c.mixed_nt[pi] = best_nt;
  c.mixed_best_ms[pi] = best_ms;
  return c.mixed_eb[pi] = best_eb;
}

template <typename T>
static void mixed_factored(Context &c, const T *eri, const T *x, T *y);

template <typename T>
static void sigma_mixed(Context &c, const T *eri, const T *x, T *y) {
  if (c.mixed_form != 0) return mixed_factored<T>(c, eri, x, y);
  if (c.B.has_ell) {
    // tuned staging width and slots per thread, then row parts
    const int pi = sizeof(T) == 4 ? 1 : 0;
    const int eb = mixed_cfg<T>(c, eri, x);
    if (eb > 0 && mixed_ell_cfg(c, eb, c.mixed_nt[pi], eri, x, y)) return check_launch("k_mixed_ell");
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
                       c.stream, nb, (int)nchunk, c.K, c.L, c.A.s_off, c.A.s_ent, c.A.str, c.B.str,
                       c.B.s_off, c.B.s_ent, eri, x, y);
  }
  check_launch("k_mixed");
}

template <typename T, int NT, int KV>
static void launch_mixed_multi(Context &c, const T *eri, const T *x, T *y, size_t lds) {
  auto fn = k_mixed_multi<T, NT, KV>;
  SBDX_HIP(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const int nslots = c.B.nslice * 64;
  const unsigned nsplit = cdiv(nslots, 1024 * NT);
  hipLaunchKernelGGL(fn, dim3((unsigned)c.na, nsplit), dim3(1024), lds, c.stream, (int)c.nb, c.K, c.L,
                     nslots, c.N, c.A.s_off, c.A.s_ent, c.A.str, c.B.str, c.B.ell, c.B.perm,
                     c.B.sl_off, c.B.sl_len, eri, x, y);
}

// KV vectors per launch with 2 slots per thread (the per-thread state grows
// with KV); false when the staging does not fit in LDS.
template <typename T, int KV>
static bool mixed_multi_dispatch(Context &c, const T *eri, const T *x, T *y) {
  if (!c.B.has_ell) return false;
  const size_t lds = sizeof(T) * ((size_t)c.K + 1 + (size_t)KV * c.nb);
  if (lds > c.lds_limit) return false;
  launch_mixed_multi<T, 2, KV>(c, eri, x, y, lds);
  check_launch("k_mixed_multi");
  return true;
}

// Mixed term of k contiguous vectors: groups of 4, then 2 vectors through
// k_mixed_multi while their rows fit in LDS, single vectors otherwise.
template <typename T>
static void sigma_mixed_block(Context &c, const T *eri, int k, const T *x, T *y) {
  const size_t N = c.N;
  const int pi = sizeof(T) == 4 ? 1 : 0;
  if (c.mixed_form != 0) {
    for (int v = 0; v < k; v++) sigma_mixed<T>(c, eri, x + v * N, y + v * N);
    return;
  }
  if (k >= 2 && c.multi_mode[pi] < 0) {
    // first block product of this precision: two vectors through the
    // multi-vector kernel vs two single-vector launches, into a scratch pair
    // (second of two runs each); the faster decides for later calls
    c.multi_mode[pi] = 0;
    const size_t lds = sizeof(T) * ((size_t)c.K + 1 + 2 * c.nb);
    if (c.B.has_ell && lds <= c.lds_limit) {
      TmpBuf tmp(sizeof(T) * N * 2);
      T *s = (T *)tmp.p;
      hipEvent_t e0, e1;
      SBDX_HIP(hipEventCreate(&e0));
      SBDX_HIP(hipEventCreate(&e1));
      float ms[2] = {0.f, 0.f};
      for (int mode = 0; mode < 2; mode++)
        for (int rep = 0; rep < 2; rep++) {
          SBDX_HIP(hipEventRecord(e0, c.stream));
          if (mode == 1) {
            mixed_multi_dispatch<T, 2>(c, eri, x, s);
          } else {
            sigma_mixed<T>(c, eri, x, s);
            sigma_mixed<T>(c, eri, x + N, s + N);
          }
          SBDX_HIP(hipEventRecord(e1, c.stream));
          SBDX_HIP(hipEventSynchronize(e1));
          SBDX_HIP(hipEventElapsedTime(&ms[mode], e0, e1));
        }
      SBDX_HIP(hipEventDestroy(e0));
      SBDX_HIP(hipEventDestroy(e1));
      c.multi_mode[pi] = ms[1] < ms[0] ? 1 : 0;
    }
  }
  int v = 0;
  while (v < k && c.multi_mode[pi] == 0) {
    sigma_mixed<T>(c, eri, x + v * N, y + v * N);
    v += 1;
  }
  while (v < k) {
    const int left = k - v;
    if (left >= 4 && mixed_multi_dispatch<T, 4>(c, eri, x + v * N, y + v * N)) {
      v += 4;
    } else if (left >= 2 && mixed_multi_dispatch<T, 2>(c, eri, x + v * N, y + v * N)) {
      v += 2;
    } else {
      sigma_mixed<T>(c, eri, x + v * N, y + v * N);
      v += 1;
    }
  }
}

// ---- same-spin terms as GEMMs
static rocblas_handle blas(Context &c) {
  if (!c.blas) {
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success)
      throw Error(3, "rocblas_create_handle failed");
    c.blas = h;
    if (rocblas_set_stream(h, c.stream) != rocblas_status_success)
      throw Error(3, "rocblas_set_stream failed");
  }
  return (rocblas_handle)c.blas;
}

void *blas_handle(Context &c) { return (void *)blas(c); }

void same_dense_release(Context &c) {
  if (c.blas) (void)rocblas_destroy_handle((rocblas_handle)c.blas);
  c.blas = nullptr;
}

template <typename T>
static T *densify(Context &c, const SpinTables &t) {
  const size_t n = (size_t)t.n;
  T *H = (T *)dalloc(c, sizeof(T) * n * n);
  SBDX_HIP(hipMemsetAsync(H, 0, sizeof(T) * n * n, c.stream));
  hipLaunchKernelGGL((k_densify<T>), dim3((unsigned)std::min<size_t>(n, 65535)), dim3(256), 0, c.stream,
                     (int)n, t.s_off, t.s_ent, t.s_val, t.d_off, t.d_ket, t.d_val, H);
  check_launch("k_densify");
  return H;
}

static void dense_blocks(Context &c, double *&Ha, double *&Hb) {
  if (!c.Hd_a) {
    c.Hd_a = densify<double>(c, c.A);
    c.Hd_b = c.B.alias ? c.Hd_a : densify<double>(c, c.B);
  }
  Ha = c.Hd_a;
  Hb = c.Hd_b;
}
static void dense_blocks(Context &c, float *&Ha, float *&Hb) {
  if (!c.Hf_a) {
    c.Hf_a = densify<float>(c, c.A);
    c.Hf_b = c.B.alias ? c.Hf_a : densify<float>(c, c.B);
  }
  Ha = c.Hf_a;
  Hb = c.Hf_b;
}

static rocblas_status gemm(rocblas_handle h, int m, int n, int k, const double *A, int lda,
                           const double *B, int ldb, double *C, int ldc, double beta = 1.0) {
  const double one = 1.0;
  return rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &one, A, lda, B,
                       ldb, &beta, C, ldc);
}
static rocblas_status gemm(rocblas_handle h, int m, int n, int k, const float *A, int lda,
                           const float *B, int ldb, float *C, int ldc, float beta = 1.0f) {
  const float one = 1.0f;
  return rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &one, A, lda, B,
                       ldb, &beta, C, ldc);
}

// y = hii*x + H_a X + X H_b.  Column-major, the na x nb row-major X is the
// nb x na matrix X^T, so with symmetric H_a, H_b:
//   (H_a X)^T = X^T H_a   -> gemm(nb, na, na, X^T, H_a)
//   (X H_b)^T = H_b X^T   -> gemm(nb, na, nb, H_b, X^T)
template <typename T>
static void same_dense(Context &c, const T *hii, const T *x, T *y) {
  T *Ha, *Hb;
  dense_blocks(c, Ha, Hb);
  const int na = (int)c.na, nb = (int)c.nb;
  const unsigned g = (unsigned)std::min<size_t>(cdiv(c.N, 256), 65536);
  hipLaunchKernelGGL((k_diag_mul<T>), dim3(g), dim3(256), 0, c.stream, c.N, hii, x, y);
  check_launch("k_diag_mul");
  rocblas_handle h = blas(c);
  if (gemm(h, nb, na, na, x, nb, Ha, na, y, nb) != rocblas_status_success)
    throw Error(3, "same-spin alpha GEMM failed");
  if (gemm(h, nb, na, nb, Hb, nb, x, nb, y, nb) != rocblas_status_success)
    throw Error(3, "same-spin beta GEMM failed");
}

static rocblas_status gemm_batched(rocblas_handle h, int m, int n, int k, const double *A, int lda,
                                   rocblas_stride sa, const double *B, int ldb, rocblas_stride sb,
                                   double *C, int ldc, rocblas_stride sc, int batch) {
  const double one = 1.0;
  return rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, m, n, k,
                                       &one, A, lda, sa, B, ldb, sb, &one, C, ldc, sc, batch);
}
static rocblas_status gemm_batched(rocblas_handle h, int m, int n, int k, const float *A, int lda,
                                   rocblas_stride sa, const float *B, int ldb, rocblas_stride sb,
                                   float *C, int ldc, rocblas_stride sc, int batch) {
  const float one = 1.0f;
  return rocblas_sgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, m, n, k,
                                       &one, A, lda, sa, B, ldb, sb, &one, C, ldc, sc, batch);
}

// Dense form for k contiguous vectors: the alpha term as a strided batch
// over the vectors (shared H_a), the beta term as one GEMM over all k*na
// columns of the side-by-side X^T blocks.
template <typename T>
static void same_dense_block(Context &c, const T *hii, int k, const T *x, T *y) {
  T *Ha, *Hb;
  dense_blocks(c, Ha, Hb);
  const int na = (int)c.na, nb = (int)c.nb;
  const size_t N = c.N;
  if ((size_t)k * na >= (1ull << 31)) throw Error(3, "block sigma: too many columns for one GEMM");
  const unsigned g = (unsigned)std::min<size_t>(cdiv(N, 256), 65536);
  for (int v = 0; v < k; v++) {
    hipLaunchKernelGGL((k_diag_mul<T>), dim3(g), dim3(256), 0, c.stream, N, hii, x + v * N, y + v * N);
    check_launch("k_diag_mul");
  }
  rocblas_handle h = blas(c);
  if (gemm_batched(h, nb, na, na, x, nb, (rocblas_stride)N, Ha, na, 0, y, nb, (rocblas_stride)N, k) !=
      rocblas_status_success)
    throw Error(3, "same-spin alpha batched GEMM failed");
  if (gemm(h, nb, k * na, nb, Hb, nb, x, nb, y, nb) != rocblas_status_success)
    throw Error(3, "same-spin beta GEMM failed");
}

template <typename T>
static void same_sparse(Context &c, const SigmaOperands<T> &o, const T *x, T *y) {
  const int na = (int)c.na, nb = (int)c.nb;
  const bool f32 = sizeof(T) == 4;
  {
    ProfScope ps(c, f32 ? "same_alpha_f32" : "same_alpha");
    sigma_same<T>(c, c.A, o.s_val_a, o.d_val_a, na, nb, o.hii, x, y, true);
  }
  {
    ProfScope ps(c, f32 ? "transpose_f32" : "transpose");
    transpose<T>(c, x, o.T1, na, nb, false);
  }
  {
    ProfScope ps(c, f32 ? "same_beta_f32" : "same_beta");
    sigma_same<T>(c, c.B, o.s_val_b, o.d_val_b, nb, na, nullptr, o.T1, o.T2, false);
  }
  {
    ProfScope ps(c, f32 ? "transpose_add_f32" : "transpose_add");
    transpose<T>(c, o.T2, y, nb, na, true);
  }
}

// Autotuner for the same-spin form: when the dense blocks fit (below a
// quarter of free device memory, 32-bit GEMM dimensions), time both forms
// on this x (second of two runs each, HIP events on the handle's stream;
// both overwrite y, which the caller's real pass overwrites again) and keep
// the faster for this precision.  SBDX_SAME_SPIN=sparse|dense overrides.
template <typename T>
static int same_mode(Context &c, const SigmaOperands<T> &o, const T *x, T *y) {
  const int pi = sizeof(T) == 4 ? 1 : 0;
  if (c.same_mode[pi] >= 0) return c.same_mode[pi];
  const char *env = std::getenv("SBDX_SAME_SPIN");
  const size_t na = c.na, nb = c.nb;
  const size_t bytes = sizeof(T) * (na * na + (c.B.alias ? 0 : nb * nb));
  size_t free_b = 0, total_b = 0;
  SBDX_HIP(hipMemGetInfo(&free_b, &total_b));
  const bool fits = na < (1u << 31) && nb < (1u << 31) && na * nb < (1ull << 31) &&
                    bytes < free_b / 4;
  if (env && std::string(env) == "sparse") return c.same_mode[pi] = 0;
  if (env && std::string(env) == "dense") {
    if (!fits) throw Error(3, "SBDX_SAME_SPIN=dense: same-spin blocks do not fit");
    return c.same_mode[pi] = 1;
  }
  if (!fits) return c.same_mode[pi] = 0;
  hipEvent_t e0, e1;
  SBDX_HIP(hipEventCreate(&e0));
  SBDX_HIP(hipEventCreate(&e1));
  float ms[2] = {0.f, 0.f};
  for (int mode = 0; mode < 2; mode++)
    for (int rep = 0; rep < 2; rep++) {
      SBDX_HIP(hipEventRecord(e0, c.stream));
      if (mode == 0)
        same_sparse<T>(c, o, x, y);
      else
        same_dense<T>(c, o.hii, x, y);
      SBDX_HIP(hipEventRecord(e1, c.stream));
      SBDX_HIP(hipEventSynchronize(e1));
      SBDX_HIP(hipEventElapsedTime(&ms[mode], e0, e1));
    }
  SBDX_HIP(hipEventDestroy(e0));
  SBDX_HIP(hipEventDestroy(e1));
  c.same_ms[pi][0] = ms[0];
  c.same_ms[pi][1] = ms[1];
  return c.same_mode[pi] = ms[1] < ms[0] ? 1 : 0;
}

template <typename T>
static void sigma_impl(Context &c, const T *x, T *y) {
  const SigmaOperands<T> o = operands(c, (T *)nullptr);
  const bool f32 = sizeof(T) == 4;
  if (same_mode<T>(c, o, x, y) == 1) {
    ProfScope ps(c, f32 ? "same_gemm_f32" : "same_gemm");
    same_dense<T>(c, o.hii, x, y);
  } else {
    same_sparse<T>(c, o, x, y);
  }
  {
    ProfScope ps(c, f32 ? "mixed_f32" : "mixed");
    sigma_mixed<T>(c, o.eri, x, y);
  }
}

void sigma_alpha_first(Context &c, const double *x, double *y) { sigma_impl<double>(c, x, y); }

// out (3n x n, column-major): column j = [hi_j; lo_j; hi_j] for the n x n
// column-major FP16 matrices hi, lo.
__global__ void __launch_bounds__(256) k_stack3(int n, const _Float16 *__restrict__ hi,
                                                const _Float16 *__restrict__ lo, _Float16 *__restrict__ out) {
  const size_t n3 = (size_t)3 * n, m = n3 * n;
  for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < m; idx += (size_t)gridDim.x * 256) {
    const size_t j = idx / n3, t3 = idx % n3, b = t3 / n, t = t3 % n;
    out[idx] = (b == 1 ? lo : hi)[t + (size_t)n * j];
  }
}

static void split_f16(Context &c, size_t n, const float *v, void *hi, void *lo, float *inv,
                      void *hi2 = nullptr) {
  const unsigned g = (unsigned)std::min<size_t>(cdiv(n, 256), 8192);
  hipLaunchKernelGGL(k_absmax_part, dim3(kSplitParts), dim3(256), 0, c.stream, n, v, c.split_part);
  check_launch("k_absmax_part");
  hipLaunchKernelGGL(k_split_f16, dim3(g), dim3(256), 0, c.stream, n, v, (const float *)c.split_part,
                     kSplitParts, (_Float16 *)hi, (_Float16 *)lo, (_Float16 *)hi2, inv);
  check_launch("k_split_f16");
}

// Symmetric FP32 same-spin term with the split-FP16x3 GEMM:
// y = hii*x + (G + G^T), G = X H_alpha as one FP16 MFMA GEMM over k = 3n,
//   G = [Xh | Xh | Xl] [Hh; Hl; Hh]
// (the three products accumulate in the matrix cores' FP32 accumulators and
// G is written once).  c.Hh holds the stacked H operand (3n x n), built
// once; c.Xh the three x blocks (n x 3n).
static void sym_same_split(Context &c, const float *hii, const float *x, float *G, float *y) {
  const int n = (int)c.na;
  const size_t nn = (size_t)n * n;
  if (!c.Hh) {
    if (!c.split_part) c.split_part = (float *)dalloc(c, sizeof(float) * kSplitParts);
    c.split_inv = (float *)dalloc(c, sizeof(float) * 2);
    c.Hh = dalloc(c, 2 * 3 * nn);
    c.Xh = dalloc(c, 2 * 3 * nn);
    float *Ha, *Hb;
    dense_blocks(c, Ha, Hb);
    TmpBuf hl(2 * 2 * nn);
    _Float16 *h = (_Float16 *)hl.p;
    split_f16(c, nn, Ha, h, h + nn, c.split_inv + 1);
    hipLaunchKernelGGL(k_stack3, dim3((unsigned)std::min<size_t>(cdiv(3 * nn, 256), 65536)), dim3(256), 0,
                       c.stream, n, (const _Float16 *)h, (const _Float16 *)(h + nn), (_Float16 *)c.Hh);
    check_launch("k_stack3");
    SBDX_HIP(hipStreamSynchronize(c.stream));  // before the halves are freed
  }
  _Float16 *X3 = (_Float16 *)c.Xh;
  split_f16(c, nn, x, X3, X3 + 2 * nn, c.split_inv, X3 + nn);
  const float one = 1.0f, zero = 0.0f;
  if (rocblas_gemm_ex(blas(c), rocblas_operation_none, rocblas_operation_none, n, n, 3 * n, &one, X3,
                      rocblas_datatype_f16_r, n, c.Hh, rocblas_datatype_f16_r, 3 * n, &zero, G,
                      rocblas_datatype_f32_r, n, G, rocblas_datatype_f32_r, n, rocblas_datatype_f32_r,
                      rocblas_gemm_algo_standard, 0, 0) != rocblas_status_success)
    throw Error(3, "split-FP16 same-spin GEMM failed");
  hipLaunchKernelGGL((k_sym_combine<float>), dim3(cdiv(n, 64), cdiv(n, 64)), dim3(256), 0, c.stream, n, hii, x,
                     (const float *)G, y, (const float *)c.split_inv, c.sym_par < 0 ? 1 : 0);
  check_launch("k_sym_combine");
}

// The symmetric same-spin term with the GEMM in the element type itself.
template <typename T>
static void sym_same_gemm(Context &c, const SigmaOperands<T> &o, const T *x, T *y) {
  const int n = (int)c.na;
  T *Ha, *Hb;
  dense_blocks(c, Ha, Hb);
  if (gemm(blas(c), n, n, n, x, n, Ha, n, o.T1, n, T(0)) != rocblas_status_success)
    throw Error(3, "symmetric same-spin GEMM failed");
  hipLaunchKernelGGL((k_sym_combine<T>), dim3(cdiv(n, 64), cdiv(n, 64)), dim3(256), 0, c.stream, n, o.hii, x,
                     (const T *)o.T1, y, (const float *)nullptr, c.sym_par < 0 ? 1 : 0);
  check_launch("k_sym_combine");
}

static void sym_same(Context &c, const SigmaOperands<double> &o, const double *x, double *y) {
  sym_same_gemm<double>(c, o, x, y);
}

// FP32: the SGEMM form or the split-FP16x3 form, decided on the first call
// by timing both after one untimed split run (which builds H's halves);
// SBDX_SPLIT_GEMM=0|1 fixes the choice.
static void sym_same(Context &c, const SigmaOperands<float> &o, const float *x, float *y) {
  if (c.split_mode < 0) {
    const char *env = std::getenv("SBDX_SPLIT_GEMM");
    if (env) {
      if (std::string(env) != "0" && std::string(env) != "1") throw Error(1, "SBDX_SPLIT_GEMM must be 0 or 1");
      c.split_mode = std::atoi(env);
    } else {
      sym_same_split(c, o.hii, x, o.T1, y);
      hipEvent_t e[3];
      for (auto &ev : e) SBDX_HIP(hipEventCreate(&ev));
      SBDX_HIP(hipEventRecord(e[0], c.stream));
      sym_same_gemm<float>(c, o, x, y);
      SBDX_HIP(hipEventRecord(e[1], c.stream));
      sym_same_split(c, o.hii, x, o.T1, y);
      SBDX_HIP(hipEventRecord(e[2], c.stream));
      SBDX_HIP(hipEventSynchronize(e[2]));
      SBDX_HIP(hipEventElapsedTime(&c.split_ms[0], e[0], e[1]));
      SBDX_HIP(hipEventElapsedTime(&c.split_ms[1], e[1], e[2]));
      for (auto &ev : e) SBDX_HIP(hipEventDestroy(ev));
      c.split_mode = c.split_ms[1] < c.split_ms[0] ? 1 : 0;
      if (c.split_mode == 1) return;  // y holds the split form's result
    }
  }
  if (c.split_mode == 1)
    sym_same_split(c, o.hii, x, o.T1, y);
  else
    sym_same_gemm<float>(c, o, x, y);
}

// ---- mixed alpha-beta term as a dense (pq|rs) contraction on the matrix cores
// The extended singles of a string are its singles (ket, pq, sign) plus, for
// each occupied orbital k, the identity entry (itself, kk, +1).  Over them the
// mixed term with both spectator dressings is
//   y[I,r] += sum_{(J,pq,sa) in ext_a(I)} sum_{(j,rs,sb) in ext_b(r)} sa sb (pq|rs) x[J,j]
//             - x[I,r] sum_{k in alpha(I)} sum_{k' in beta(r)} (kk|k'k')
// (identity-identity pairs add the diagonal alpha-beta Coulomb term, which
// hii already holds; the correction removes it).  For a batch of B beta bra
// rows r0.. and column index col = (r - r0) na + J (M = B na columns):
//   D[rs][col] = sum_{(j,rs,sb) in ext_b(r)} sb x[J,j]    K x M, mostly zero
//   G = E D with E[pq][rs] = (pq|rs)                      2 K^2 M flops, one GEMM
//   y[I,r] += sum_{(J,pq,sa) in ext_a(I)} sa G[pq][col(r,J)] - correction
// Per vector that is 2 K^2 na nb flops whatever the sparsity of D.
constexpr size_t kFacBytes = size_t(1) << 30;  // D and G, each (FP64 sizing)

template <typename T>
__global__ void __launch_bounds__(256) k_fac_zero(size_t n, T *__restrict__ v) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) v[i] = T(0);
}

// grid (ceil(na/256), B): thread J of batch row b writes column (b, J) of D.
// Within one bra r the singles' rs are distinct and off the (kk) diagonal, so
// every element is written at most once.  xT = x^T (nb x na).
// SPLIT: instead of D (n = M K entries), the FP16 operand [Dh | Dh | Dl] of
// the split GEMM into D3 (3n halves, zero-filled), with the power-of-two
// scale of k_split_f16 taken from the block maxima `part` of |x| (D holds
// +-x values or zeros); block (0, 0) stores the inverse scale at *inv.
template <typename T, bool SPLIT>
__global__ void __launch_bounds__(256) k_fac_build(int na, int L, int r0, size_t M, size_t n,
                                                   const uint64_t *__restrict__ b_off,
                                                   const uint64_t *__restrict__ b_ent,
                                                   const uint64_t *__restrict__ b_str,
                                                   const T *__restrict__ xT, T *__restrict__ D,
                                                   _Float16 *__restrict__ D3, const float *__restrict__ part,
                                                   float *__restrict__ inv) {
  float sc = 1.f;
  if (SPLIT) {  // every thread of the block takes part, before any return
    float m = 0.f;
    for (int i = threadIdx.x; i < kSplitParts; i += 256) m = fmaxf(m, part[i]);
    m = block_max256(m);
    int k = 0;
    if (m > 0.f) (void)frexpf(m, &k);  // m = f 2^k, f in [0.5, 1)
    const int e = m > 0.f ? std::min(15 - k, 100) : 0;
    sc = ldexpf(1.f, e);
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *inv = ldexpf(1.f, -e);
  }
  auto put = [&](size_t idx, T v) {
    if (SPLIT) {
      const float t = (float)v * sc;
      const _Float16 h = (_Float16)t;
      D3[idx] = h;
      D3[n + idx] = h;
      D3[2 * n + idx] = (_Float16)(t - (float)h);
    } else {
      D[idx] = v;
    }
  };
  const int J = blockIdx.x * 256 + threadIdx.x;
  if (J >= na) return;
  const int b = blockIdx.y, r = r0 + b;
  const size_t col = (size_t)b * na + J;
  for (uint64_t f = b_off[r]; f < b_off[r + 1]; f++) {
    const uint64_t be = b_ent[f];
    const T v = xT[(size_t)ent_ket(be) * na + J];
    put((size_t)ent_pq(be) * M + col, ent_neg(be) ? -v : v);
  }
  const T xr = xT[(size_t)r * na + J];
  for (uint64_t occ = b_str[r]; occ; occ &= occ - 1) put((size_t)(__builtin_ctzll(occ) * (L + 1)) * M + col, xr);
}

// thread (I, b), b fastest: y[I, r0+b] += the alpha side over G.  gscale
// (device, may be null): G holds 2^(ex+eh) times the product.
template <typename T>
__global__ void __launch_bounds__(256) k_fac_apply(int na, int nb, int L, int r0, int B, size_t M,
                                                   const uint64_t *__restrict__ a_off,
                                                   const uint64_t *__restrict__ a_ent,
                                                   const uint64_t *__restrict__ a_str,
                                                   const double *__restrict__ jsum_b,
                                                   const T *__restrict__ G, const float *__restrict__ gscale,
                                                   const T *__restrict__ x, T *__restrict__ y) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (size_t)na * B) return;
  const int b = (int)(t % B), I = (int)(t / B), r = r0 + b;
  const size_t cb = (size_t)b * na;
  T g = T(0);
  for (uint64_t e = a_off[I]; e < a_off[I + 1]; e++) {
    const uint64_t ae = a_ent[e];
    const T v = G[(size_t)ent_pq(ae) * M + cb + ent_ket(ae)];
    g += ent_neg(ae) ? -v : v;
  }
  double corr = 0.0;
  for (uint64_t occ = a_str[I]; occ; occ &= occ - 1) {
    const int k = __builtin_ctzll(occ);
    g += G[(size_t)(k * (L + 1)) * M + cb + I];
    corr += jsum_b[(size_t)r * L + k];  // sum_{k' in beta(r)} (k'k'|kk)
  }
  if (gscale) g *= (T)(gscale[0] * gscale[1]);
  const size_t o = (size_t)I * nb + r;
  y[o] += g - x[o] * (T)corr;
}

// G and x^T on first use; D (in the product's own precision) only for the
// DGEMM/SGEMM forms, since the split form writes its FP16 operand directly
static void fac_workspace(Context &c, bool need_D) {
  if (!c.fac_G) {
    const size_t row = sizeof(double) * (size_t)c.K * c.na;  // one beta row of D
    c.fac_cap = std::max(row, std::min(kFacBytes, row * c.nb));
    c.fac_G = dalloc(c, c.fac_cap);
    c.fac_xT = dalloc(c, sizeof(double) * c.N);
  }
  if (need_D && !c.fac_D) c.fac_D = dalloc(c, c.fac_cap);
}

// G = E D for the batch: the product's own precision (DGEMM / SGEMM), or for
// FP32 with form 2 the split-FP16x3 GEMM (D and E as 2^-e (hi + lo), see
// k_split_f16) whose scales are returned for the apply kernel.  The three
// FP16 products run as one GEMM over k = 3K,
//   G = [Dh | Dh | Dl] [Eh; El; Eh],
// so they accumulate in the matrix cores' FP32 accumulators and G is
// written once.  The ERI operand is stacked per column (3K x K, column pq =
// [Eh; El; Eh] of ERI row pq), built once.
static const float *fac_gemm(Context &c, int M, const double *eri, const double *D, double *G) {
  if (gemm(blas(c), M, c.K, c.K, D, M, eri, c.K, G, M, 0.0) != rocblas_status_success)
    throw Error(3, "factored mixed-term DGEMM failed");
  return nullptr;
}
// FP16 operands of the split form, built on first use: the ERI matrix
// stacked per column ([Eh; El; Eh], 3K x K), room for [Dh | Dh | Dl] of an
// FP32 D of at most fac_cap/4 entries, the block maxima of |x| and the two
// inverse scales (D's, ERI's).
static void fac_split_init(Context &c, const float *eri) {
  if (c.fac_E3) return;
  const size_t KK = (size_t)c.K * c.K;
  if (!c.split_part) c.split_part = (float *)dalloc(c, sizeof(float) * kSplitParts);
  c.fac_inv = (float *)dalloc(c, sizeof(float) * 2);
  c.fac_part = (float *)dalloc(c, sizeof(float) * kSplitParts);
  c.fac_D3 = dalloc(c, 3 * 2 * (c.fac_cap / 4));
  _Float16 *E3 = (_Float16 *)dalloc(c, 2 * 3 * KK);
  TmpBuf hl(2 * 2 * KK);
  _Float16 *h = (_Float16 *)hl.p;
  split_f16(c, KK, eri, h, h + KK, c.fac_inv + 1);
  hipLaunchKernelGGL(k_stack3, dim3((unsigned)std::min<size_t>(cdiv(3 * KK, 256), 65536)), dim3(256), 0, c.stream,
                     c.K, (const _Float16 *)h, (const _Float16 *)(h + KK), E3);
  check_launch("k_stack3");
  SBDX_HIP(hipStreamSynchronize(c.stream));  // before the halves are freed
  c.fac_E3 = E3;
}

static const float *fac_gemm(Context &c, int M, const float *eri, const float *D, float *G) {
  rocblas_handle hb = blas(c);
  if (c.mixed_form != 2) {
    if (gemm(hb, M, c.K, c.K, D, M, eri, c.K, G, M, 0.f) != rocblas_status_success)
      throw Error(3, "factored mixed-term SGEMM failed");
    return nullptr;
  }
  // [Dh | Dh | Dl] was written by k_fac_build<float, true>
  const int K = c.K;
  const float one = 1.0f, zero = 0.0f;
  if (rocblas_gemm_ex(hb, rocblas_operation_none, rocblas_operation_none, M, K, 3 * K, &one, c.fac_D3,
                      rocblas_datatype_f16_r, M, c.fac_E3, rocblas_datatype_f16_r, 3 * K, &zero, G,
                      rocblas_datatype_f32_r, M, G, rocblas_datatype_f32_r, M, rocblas_datatype_f32_r,
                      rocblas_gemm_algo_standard, 0, 0) != rocblas_status_success)
    throw Error(3, "factored mixed-term split-FP16 GEMM failed");
  return c.fac_inv;
}

template <typename T>
static void mixed_factored(Context &c, const T *eri, const T *x, T *y) {
  fac_workspace(c, !(sizeof(T) == 4 && c.mixed_form == 2));
  const int na = (int)c.na, nb = (int)c.nb, K = c.K;
  const size_t row = (size_t)K * na;
  // beta rows per batch: as many as fit in the workspace, at most 32768
  // (a grid dimension), or SBDX_FACTORED_ROWS when set
  size_t rows = std::min<size_t>({(size_t)nb, c.fac_cap / (sizeof(T) * row), 32768});
  if (const char *env = std::getenv("SBDX_FACTORED_ROWS")) {
    const long v = std::atol(env);
    if (v < 1) throw Error(1, "SBDX_FACTORED_ROWS must be a positive integer");
    rows = std::min<size_t>(rows, (size_t)v);
  }
  const int B = (int)std::max<size_t>(1, rows);
  if ((size_t)B * na > (size_t)INT32_MAX) throw Error(3, "factored mixed term: batch too wide for the GEMM");
  T *xT = (T *)c.fac_xT, *D = (T *)c.fac_D, *G = (T *)c.fac_G;
  transpose<T>(c, x, xT, na, nb, false);
  const bool f32 = sizeof(T) == 4;
  bool split = false;
  if constexpr (sizeof(T) == 4) {
    if (c.mixed_form == 2) {  // FP16 operands, and the block maxima of |x| for D's scale
      split = true;
      fac_split_init(c, eri);
      hipLaunchKernelGGL(k_absmax_part, dim3(kSplitParts), dim3(256), 0, c.stream, c.N, x, c.fac_part);
      check_launch("k_absmax_part");
    }
  }
  for (int r0 = 0; r0 < nb; r0 += B) {
    const int Bc = std::min(B, nb - r0);
    const size_t M = (size_t)Bc * na, n = M * K;
    const dim3 gb(cdiv(na, 256), (unsigned)Bc);
    if (split) {
      _Float16 *D3 = (_Float16 *)c.fac_D3;
      hipLaunchKernelGGL((k_fac_zero<_Float16>), dim3((unsigned)std::min<size_t>(cdiv(3 * n, 256), 65536)),
                         dim3(256), 0, c.stream, 3 * n, D3);
      check_launch("k_fac_zero");
      hipLaunchKernelGGL((k_fac_build<T, true>), gb, dim3(256), 0, c.stream, na, c.L, r0, M, n, c.B.s_off,
                         c.B.s_ent, c.B.str, (const T *)xT, D, D3, (const float *)c.fac_part, c.fac_inv);
    } else {
      hipLaunchKernelGGL((k_fac_zero<T>), dim3((unsigned)std::min<size_t>(cdiv(n, 256), 65536)), dim3(256), 0,
                         c.stream, n, D);
      check_launch("k_fac_zero");
      hipLaunchKernelGGL((k_fac_build<T, false>), gb, dim3(256), 0, c.stream, na, c.L, r0, M, n, c.B.s_off,
                         c.B.s_ent, c.B.str, (const T *)xT, D, (_Float16 *)nullptr, (const float *)nullptr,
                         (float *)nullptr);
    }
    check_launch("k_fac_build");
    const float *gs;
    {
      ProfScope ps(c, f32 ? "fac_gemm_f32" : "fac_gemm");
      gs = fac_gemm(c, (int)M, eri, D, G);
    }
    hipLaunchKernelGGL((k_fac_apply<T>), dim3(cdiv(M, 256)), dim3(256), 0, c.stream, na, nb, c.L, r0, Bc, M,
                       c.A.s_off, c.A.s_ent, c.A.str, (const double *)c.B.jsum, (const T *)G, gs, x, y);
    check_launch("k_fac_apply");
  }
}

// Whether the symmetric form applies: equal string sets, the dense blocks
// fit (same-spin autotuner chose or can hold them), the sliced-ELL table.
static bool sym_ok(Context &c) {
  if (!c.B.alias || !c.B.has_ell || c.na != c.nb) return false;
  size_t free_b = 0, total_b = 0;
  SBDX_HIP(hipMemGetInfo(&free_b, &total_b));
  const size_t bytes = (c.Hd_a ? 0 : sizeof(double) * c.na * c.na) +
                       (c.f32.ready && !c.Hf_a ? sizeof(float) * c.na * c.na : 0);
  return c.na < (1u << 31) && c.N < (1ull << 31) && bytes < free_b / 4;
}

// sigma for a spin-flip symmetric x (x = x^T as an na x na matrix, equal
// string sets): one GEMM G = H_a X, y = hii*x + G + G^T, and the mixed term
// for rows r >= I only, mirrored.  The result equals the general sigma for
// such x (H commutes with the transpose when the string sets are equal).
template <typename T>
static void sigma_sym_impl(Context &c, const T *x, T *y) {
  if (c.mixed_form != 0) {
    // the factored mixed term has no triangle form: the general product,
    // equal for symmetric x
    sigma_impl<T>(c, x, y);
    return;
  }
  const SigmaOperands<T> o = operands(c, (T *)nullptr);
  const bool f32 = sizeof(T) == 4;
  {
    ProfScope ps(c, f32 ? "sym_same_f32" : "sym_same");
    sym_same(c, o, x, y);
  }
  if (sizeof(T) * ((size_t)c.nb + c.K + 1) > c.lds_limit) {
    // one staged pair does not fit in LDS (e.g. FP64 at 20000 strings): the
    // ket rows staged in parts, same triangle of slots; if even that does not
    // fit, the general product (equal for symmetric x)
    ProfScope ps(c, f32 ? "sym_mixed_f32" : "sym_mixed");
    if (!mixed_parts_dispatch<T, true>(c, o.eri, x, y)) {
      sigma_impl<T>(c, x, y);
      return;
    }
    check_launch("k_mixed_parts<sym>");
    return;
  }
  {
    ProfScope ps(c, f32 ? "sym_mixed_f32" : "sym_mixed");
    const int pi = (f32 ? 1 : 0) + 2;
    const int eb = mixed_cfg<T, true>(c, o.eri, x);
    const int nt = c.mixed_nt[pi];
    bool ok = false;
    if (eb == kEbSigned) ok = mixed_sw_dispatch<T, true>(c, nt, o.eri, x, y);
    else if (eb == 4) ok = mixed_ell_dispatch<T, 4, true>(c, nt, o.eri, x, y);
    else if (eb == 2) ok = mixed_ell_dispatch<T, 2, true>(c, nt, o.eri, x, y);
    else if (eb == 1) ok = mixed_ell_dispatch<T, 1, true>(c, nt, o.eri, x, y);
    if (!ok) throw Error(3, "symmetric mixed kernel: staging does not fit");
    check_launch("k_mixed_ell<sym>");
  }
}

bool sigma_sym_available(Context &c) { return sym_ok(c); }

void sigma_sym(Context &c, const double *x, double *y) { sigma_sym_impl<double>(c, x, y); }

// The FP32 symmetric core (same-spin GEMM + combine + mixed kernel on the
// fixed FP32 staging buffers) as a replayed hipGraph: the first call runs
// eagerly (autotuning, dense blocks and rocBLAS workspace are settled then)
// and captures the same sequence for later calls.
static void sym_core_f32(Context &c) {
  if (!c.use_graphs || c.mixed_form != 0 || sizeof(float) * ((size_t)c.nb + c.K + 1) > c.lds_limit) {
    sigma_sym_impl<float>(c, c.f32.x, c.f32.y);
    return;
  }
  if (c.g_sym32 && c.g_sym32_par != c.sym_par) graphs_release(c);  // captured for the other sector
  if (c.g_sym32) {
    ProfScope ps(c, "sym_graph_f32");
    SBDX_HIP(hipGraphLaunch(c.g_sym32, c.stream));
    return;
  }
  sigma_sym_impl<float>(c, c.f32.x, c.f32.y);  // eager; also this call's result
  SBDX_HIP(hipStreamSynchronize(c.stream));
  hipGraph_t graph = nullptr;
  c.capturing = true;
  SBDX_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeRelaxed));
  try {
    sigma_sym_impl<float>(c, c.f32.x, c.f32.y);
  } catch (...) {
    (void)hipStreamEndCapture(c.stream, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    c.capturing = false;
    throw;
  }
  c.capturing = false;
  SBDX_HIP(hipStreamEndCapture(c.stream, &graph));
  SBDX_HIP(hipGraphInstantiate(&c.g_sym32, graph, nullptr, nullptr, 0));
  SBDX_HIP(hipGraphDestroy(graph));
  c.g_sym32_par = c.sym_par;
}

void graphs_release(Context &c) {
  if (c.g_sym32) (void)hipGraphExecDestroy(c.g_sym32);
  c.g_sym32 = nullptr;
}

void sigma_sym_f32_raw(Context &c, const float *x, float *y) {
  ensure_f32(c);
  sigma_sym_impl<float>(c, x, y);
}

// The FP32 symmetric product of an FP64 vector, left in FP32 (the values
// sigma_sym_f32 widens).
void sigma_sym_f32_to(Context &c, const double *x, float *y) {
  ensure_f32(c);
  const unsigned g = (unsigned)std::min<size_t>(cdiv(c.N, 256), 65536);
  hipLaunchKernelGGL((k_cast<double, float>), dim3(g), dim3(256), 0, c.stream, c.N, x, c.f32.x);
  check_launch("k_cast");
  sym_core_f32(c);
  SBDX_HIP(hipMemcpyAsync(y, c.f32.y, sizeof(float) * c.N, hipMemcpyDeviceToDevice, c.stream));
}

void sigma_sym_f32(Context &c, const double *x, double *y) {
  ensure_f32(c);
  const unsigned g = (unsigned)std::min<size_t>(cdiv(c.N, 256), 65536);
  hipLaunchKernelGGL((k_cast<double, float>), dim3(g), dim3(256), 0, c.stream, c.N, x, c.f32.x);
  check_launch("k_cast");
  sym_core_f32(c);
  hipLaunchKernelGGL((k_cast<float, double>), dim3(g), dim3(256), 0, c.stream, c.N, c.f32.y, y);
  check_launch("k_cast");
}

// max |v - p v^T| / max |v| of an na x na vector (0 for v = 0); p = +1
// measures the distance from the even sector, p = -1 from the odd one
double sym_defect(Context &c, const double *v, int p) {
  const int n = (int)c.na;
  const unsigned t = cdiv(n, 64);
  TmpBuf part(sizeof(double) * 2 * t * t);
  hipLaunchKernelGGL(k_sym_defect, dim3(t, t), dim3(256), 0, c.stream, n, v, p < 0 ? -1.0 : 1.0,
                     (double *)part.p);
  check_launch("k_sym_defect");
  std::vector<double> h((size_t)2 * t * t);
  SBDX_HIP(hipMemcpyAsync(h.data(), part.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c.stream));
  SBDX_HIP(hipStreamSynchronize(c.stream));
  double d = 0.0, m = 0.0;
  for (size_t b = 0; b < (size_t)t * t; b++) {
    d = std::max(d, h[2 * b]);
    m = std::max(m, h[2 * b + 1]);
  }
  return m > 0.0 ? d / m : 0.0;
}

void symmetrize(Context &c, double *v, int p) {
  const unsigned t = cdiv(c.na, 64);
  hipLaunchKernelGGL(k_symmetrize, dim3(t, t), dim3(256), 0, c.stream, (int)c.na, v, p < 0 ? -1.0 : 1.0);
  check_launch("k_symmetrize");
}

// k products at once, vectors contiguous (x + v*N -> y + v*N).
template <typename T>
static void sigma_block_impl(Context &c, int k, const T *x, T *y) {
  const SigmaOperands<T> o = operands(c, (T *)nullptr);
  const bool f32 = sizeof(T) == 4;
  const size_t N = c.N;
  if (same_mode<T>(c, o, x, y) == 1) {
    ProfScope ps(c, f32 ? "blk_same_gemm_f32" : "blk_same_gemm");
    same_dense_block<T>(c, o.hii, k, x, y);
  } else {
    for (int v = 0; v < k; v++) same_sparse<T>(c, o, x + v * N, y + v * N);
  }
  {
    ProfScope ps(c, f32 ? "blk_mixed_f32" : "blk_mixed");
    sigma_mixed_block<T>(c, o.eri, k, x, y);
  }
}

void sigma_block(Context &c, int k, const double *x, double *y) {
  if (k < 1) return;
  sigma_block_impl<double>(c, k, x, y);
}

void sigma_block_f32(Context &c, int k, const double *x, double *y) {
  if (k < 1) return;
  ensure_f32(c);
  const size_t n = (size_t)k * c.N;
  if (c.f32.kcap < k) {
    c.f32.xb = (float *)dalloc(c, sizeof(float) * n);
    c.f32.yb = (float *)dalloc(c, sizeof(float) * n);
    c.f32.kcap = k;
  }
  const unsigned g = (unsigned)std::min<size_t>(cdiv(n, 256), 65536);
  hipLaunchKernelGGL((k_cast<double, float>), dim3(g), dim3(256), 0, c.stream, n, x, c.f32.xb);
  check_launch("k_cast");
  sigma_block_impl<float>(c, k, c.f32.xb, c.f32.yb);
  hipLaunchKernelGGL((k_cast<float, double>), dim3(g), dim3(256), 0, c.stream, n, c.f32.yb, y);
  check_launch("k_cast");
}

void sigma_block_f32_raw(Context &c, int k, const float *x, float *y) {
  if (k < 1) return;
  ensure_f32(c);
  sigma_block_impl<float>(c, k, x, y);
}

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
  if (c.B.alias) {
    c.f32.s_val_b = c.f32.s_val_a;
    c.f32.d_val_b = c.f32.d_val_a;
  } else {
    c.f32.s_val_b = cast(c.B.s_val, c.B.ns);
    c.f32.d_val_b = cast(c.B.d_val, c.B.nd);
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
