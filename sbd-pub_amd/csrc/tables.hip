// Excitation tables, the diagonal and the basis sign.
//
// Table construction follows the semantics of GenerateExcitation
// (chemistry/tpb/helper.h:185-323): for every bra string, every ket string of
// the same set at Hamming distance 2 (single) or 4 (double), in ascending ket
// index.  One 64-lane wavefront owns one bra string and sweeps the kets 64 at
// a time; popcount of the XOR gives the distance and the wave ballot + a
// masked popcount give each match its slot, so the order is ascending
// without sorting.
#include <algorithm>
#include <array>
#include <atomic>
#include <thread>

#include "common.h"

namespace sbdx {

__device__ __forceinline__ double parity1(uint64_t s, int a, int b) {
  // Reference parity rule (chemistry/basic/determinants.h:230-281) on one
  // string: bits in [lo, hi), one more flip when bit lo is set.
  int lo = a < b ? a : b, hi = a < b ? b : a;
  uint64_t mask = ((1ull << hi) - 1ull) ^ ((1ull << lo) - 1ull);
  int n = __popcll(s & mask) + (int)((s >> lo) & 1ull);
  return (n & 1) ? -1.0 : 1.0;
}

__global__ void __launch_bounds__(256) k_table_count(const uint64_t *__restrict__ s, int n,
                                                     uint32_t *__restrict__ cs,
                                                     uint32_t *__restrict__ cd) {
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (wave >= n) return;  // uniform per wave
  const uint64_t me = s[wave];
  uint32_t a = 0, b = 0;
  for (int j0 = 0; j0 < n; j0 += 64) {
    const int j = j0 + lane;
    const int d = (j < n) ? __popcll(me ^ s[j]) : 0;
    a += (uint32_t)__popcll(__ballot(d == 2));
    b += (uint32_t)__popcll(__ballot(d == 4));
  }
  if (lane == 0) {
    cs[wave] = a;
    cd[wave] = b;
  }
}

__global__ void __launch_bounds__(256) k_table_fill(
    const uint64_t *__restrict__ s, int n, int L, const uint64_t *__restrict__ s_off,
    const uint64_t *__restrict__ d_off, const double *__restrict__ eri,
    const double *__restrict__ h1, uint64_t *__restrict__ s_ent, double *__restrict__ s_val,
    uint32_t *__restrict__ d_ket, uint32_t *__restrict__ d_orb, double *__restrict__ d_val) {
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (wave >= n) return;
  const int K = L * L;
  const uint64_t me = s[wave];
  const uint64_t below = (1ull << lane) - 1ull;
  uint64_t ps = s_off[wave], pd = d_off[wave];
  for (int j0 = 0; j0 < n; j0 += 64) {
    const int j = j0 + lane;
    const uint64_t sj = (j < n) ? s[j] : me;
    const int d = __popcll(me ^ sj);
    const uint64_t m2 = __ballot(d == 2), m4 = __ballot(d == 4);
    if (d == 2) {
      const uint64_t pos = ps + (uint64_t)__popcll(m2 & below);
      const int p = __ffsll((unsigned long long)(me & ~sj)) - 1;
      const int q = __ffsll((unsigned long long)(sj & ~me)) - 1;
      const double sg = parity1(me, p, q);
      const int pq = p * L + q;
      double F = h1[pq];
      uint64_t occ = me;
      while (occ) {
        const int k = __ffsll((unsigned long long)occ) - 1;
        occ &= occ - 1;
        F += eri[(size_t)pq * K + k * L + k] - eri[(size_t)(p * L + k) * K + k * L + q];
      }
      s_ent[pos] = (uint64_t)j | ((uint64_t)pq << 32) | (sg < 0 ? (1ull << 63) : 0ull);
      s_val[pos] = sg * F;
    } else if (d == 4) {
      const uint64_t pos = pd + (uint64_t)__popcll(m4 & below);
      const uint64_t cr = me & ~sj, an = sj & ~me;
      const int i = __ffsll((unsigned long long)cr) - 1, jj = 63 - __clzll(cr);
      const int a = __ffsll((unsigned long long)an) - 1, b = 63 - __clzll(an);
      double sg = parity1(me, i, a) * parity1(me, jj, b);
      if (a > jj || b < i) sg = -sg;
      const double v = eri[(size_t)(a * L + i) * K + b * L + jj] -
                       eri[(size_t)(a * L + jj) * K + b * L + i];
      d_ket[pos] = (uint32_t)j;
      d_orb[pos] = (uint32_t)i | ((uint32_t)jj << 8) | ((uint32_t)a << 16) | ((uint32_t)b << 24);
      d_val[pos] = sg * v;
    }
    ps += (uint64_t)__popcll(m2);
    pd += (uint64_t)__popcll(m4);
  }
}

// Per-string pieces of the diagonal: e1[x] = sum_p h_pp + sum_{p<q}
// ((pp|qq) - (pq|qp)) and jsum[x][r] = sum_{p in x} (pp|rr).
__global__ void __launch_bounds__(256) k_string_energy(const uint64_t *__restrict__ s, int n,
                                                       int L, const double *__restrict__ eri,
                                                       const double *__restrict__ h1,
                                                       double *__restrict__ e1,
                                                       double *__restrict__ jsum) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const int K = L * L;
  const uint64_t str = s[x];
  double e = 0.0;
  for (int p = 0; p < L; p++) {
    if (!((str >> p) & 1ull)) continue;
    e += h1[p * L + p];
    for (int q = p + 1; q < L; q++) {
      if (!((str >> q) & 1ull)) continue;
      e += eri[(size_t)(p * L + p) * K + q * L + q] - eri[(size_t)(p * L + q) * K + q * L + p];
    }
  }
  e1[x] = e;
  for (int r = 0; r < L; r++) {
    double j = 0.0;
    for (int p = 0; p < L; p++)
      if ((str >> p) & 1ull) j += eri[(size_t)(p * L + p) * K + r * L + r];
    jsum[(size_t)x * L + r] = j;
  }
}

// hii[ia*nb+ib] = I0 + e1a[ia] + e1b[ib] + sum_{r in beta(ib)} jsum_a[ia][r]
// (ZeroExcite, chemistry/basic/determinants.h:434-457, split by spin).
// canon != 0 (equal string sets, shared tables): element (ia, ib) with
// ib < ia is evaluated as (ib, ia), so hii is bitwise symmetric and the
// spin-flip sectors (x = +-x^T) stay exact under hii * x.
__global__ void __launch_bounds__(256) k_diag(int nb, int nchunk, int L, double I0, int canon,
                                              const uint64_t *__restrict__ bstr,
                                              const double *__restrict__ e1a,
                                              const double *__restrict__ e1b,
                                              const double *__restrict__ jsa,
                                              double *__restrict__ hii) {
  const int ia = blockIdx.x / nchunk;
  const int ib = (blockIdx.x % nchunk) * blockDim.x + threadIdx.x;
  if (ib >= nb) return;
  const bool swap = canon && ib < ia;
  const double *j = jsa + (size_t)(swap ? ib : ia) * L;
  uint64_t occ = bstr[swap ? ia : ib];
  double acc = 0.0;
  while (occ) {
    const int r = __ffsll((unsigned long long)occ) - 1;
    occ &= occ - 1;
    acc += j[r];
  }
  hii[(size_t)ia * nb + ib] = I0 + (e1a[ia] + e1b[ib]) + acc;
}

// out = S * in, S(ia,ib) = (-1)^#{(p in alpha, r in beta): r < p}: the sign
// between the interleaved (reference) and the alpha-first basis.
template <bool ACC>
__global__ void __launch_bounds__(256) k_sign(int nb, int nchunk, const uint64_t *__restrict__ astr,
                                              const uint64_t *__restrict__ bstr,
                                              const double *__restrict__ in,
                                              double *__restrict__ out) {
  const int ia = blockIdx.x / nchunk;
  const int ib = (blockIdx.x % nchunk) * blockDim.x + threadIdx.x;
  if (ib >= nb) return;
  uint64_t a = astr[ia];
  const uint64_t b = bstr[ib];
  int cnt = 0;
  while (a) {
    const int p = __ffsll((unsigned long long)a) - 1;
    a &= a - 1;
    cnt += __popcll(b & ((1ull << p) - 1ull));
  }
  const size_t idx = (size_t)ia * nb + ib;
  const double v = (cnt & 1) ? -in[idx] : in[idx];
  out[idx] = ACC ? out[idx] + v : v;
}

static inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

void build_spin_tables(Context &c, SpinTables &t, const std::vector<uint64_t> &strs) {
  const int n = (int)strs.size();
  t.n = n;
  t.str = (uint64_t *)dalloc(c, sizeof(uint64_t) * n);
  SBDX_HIP(hipMemcpyAsync(t.str, strs.data(), sizeof(uint64_t) * n, hipMemcpyHostToDevice, c.stream));
  uint32_t *cnt = (uint32_t *)dalloc(c, sizeof(uint32_t) * 2 * (size_t)n);
  {
    ProfScope ps(c, "table_count");
    hipLaunchKernelGGL(k_table_count, dim3(cdiv((size_t)n * 64, 256)), dim3(256), 0, c.stream,
                       t.str, n, cnt, cnt + n);
    check_launch("k_table_count");
  }
  std::vector<uint32_t> hc(2 * (size_t)n);
  SBDX_HIP(hipMemcpyAsync(hc.data(), cnt, sizeof(uint32_t) * 2 * n, hipMemcpyDeviceToHost, c.stream));
  SBDX_HIP(hipStreamSynchronize(c.stream));
  std::vector<uint64_t> so(n + 1, 0), dof(n + 1, 0);
  size_t mx = 0;
  for (int i = 0; i < n; i++) {
    so[i + 1] = so[i] + hc[i];
    dof[i + 1] = dof[i] + hc[n + i];
    if (hc[i] > mx) mx = hc[i];
  }
  t.ns = so[n];
  t.nd = dof[n];
  t.max_s = mx;
  t.s_off = (uint64_t *)dalloc(c, sizeof(uint64_t) * (n + 1));
  t.d_off = (uint64_t *)dalloc(c, sizeof(uint64_t) * (n + 1));
  SBDX_HIP(hipMemcpyAsync(t.s_off, so.data(), sizeof(uint64_t) * (n + 1), hipMemcpyHostToDevice, c.stream));
  SBDX_HIP(hipMemcpyAsync(t.d_off, dof.data(), sizeof(uint64_t) * (n + 1), hipMemcpyHostToDevice, c.stream));
  t.s_ent = (uint64_t *)dalloc(c, sizeof(uint64_t) * (t.ns + 1));
  t.s_val = (double *)dalloc(c, sizeof(double) * (t.ns + 1));
  t.d_ket = (uint32_t *)dalloc(c, sizeof(uint32_t) * (t.nd + 1));
  t.d_orb = (uint32_t *)dalloc(c, sizeof(uint32_t) * (t.nd + 1));
  t.d_val = (double *)dalloc(c, sizeof(double) * (t.nd + 1));
  {
    ProfScope ps(c, "table_fill");
    hipLaunchKernelGGL(k_table_fill, dim3(cdiv((size_t)n * 64, 256)), dim3(256), 0, c.stream, t.str,
                       n, c.L, t.s_off, t.d_off, c.eri, c.h1, t.s_ent, t.s_val, t.d_ket, t.d_orb,
                       t.d_val);
    check_launch("k_table_fill");
  }
  t.e1 = (double *)dalloc(c, sizeof(double) * n);
  t.jsum = (double *)dalloc(c, sizeof(double) * (size_t)n * c.L);
  hipLaunchKernelGGL(k_string_energy, dim3(cdiv(n, 256)), dim3(256), 0, c.stream, t.str, n, c.L,
                     c.eri, c.h1, t.e1, t.jsum);
  check_launch("k_string_energy");
}

// Sliced-ELL copy of the singles (see SpinTables::ell).  Needs kets < 2^16
// and pq < 2^15; otherwise has_ell stays false and the CSR kernels run.
// Order of one 64-row slice's entries (position k of lane l at
// out[64*k + l]).  The mixed kernels read x[.., ket] and V[.., pq] from LDS
// at the indices of entry k of all 64 lanes at once, and the row sums do
// not depend on the entry order, so entries are dealt greedily: at each k,
// lane by lane, the remaining entry whose ket and pq residues are least
// used so far at this k.  Measured gfx950 LDS bank rule: 4- and 8-byte
// reads are served 32 lanes at a time and two distinct addresses of one
// half-wavefront conflict when their element indices agree mod 32; 16-byte
// reads (4 FP32 or 2 FP64 pairs per entry) are served 16 lanes at a time and
// conflict when the indices agree mod 16 within the quarter-wavefront.  The
// cost of a candidate is 2 x (uses of its residues mod 16 in the lane's
// quarter) + (uses of its residues mod 32 in the lane's half), which keeps
// both rules nearly balanced.  Padding entries (pq = pad, a zero in every
// staged ERI row, one shared address) take an unused ket residue below nb.
static void schedule_slice(std::vector<uint32_t> *rows, uint32_t len, uint32_t pad_pq, uint32_t nb,
                           uint32_t *out) {
  for (uint32_t k = 0; k < len; k++) {
    uint8_t hj[32], hv[32], qj[16], qv[16];
    for (int l = 0; l < 64; l++) {
      if ((l & 31) == 0) {
        std::fill(hj, hj + 32, 0);
        std::fill(hv, hv + 32, 0);
      }
      if ((l & 15) == 0) {
        std::fill(qj, qj + 16, 0);
        std::fill(qv, qv + 16, 0);
      }
      std::vector<uint32_t> &R = rows[l];
      uint32_t w;
      if (R.empty()) {
        int best = 0, bc = 1 << 30;
        for (int j = 0; j < 32; j++) {
          const int cost = 2 * qj[j & 15] + hj[j];
          if (cost < bc) {
            bc = cost;
            best = j;
          }
        }
        const uint32_t jb = nb > 32 ? (uint32_t)best : (uint32_t)best % nb;
        w = jb | (pad_pq << 16);
      } else {
        size_t bi = 0;
        int bc = 1 << 30;
        for (size_t i = 0; i < R.size(); i++) {
          const uint32_t e = R[i], v = e >> 16;
          const int cost = 2 * (qj[e & 15] + qv[v & 15]) + hj[e & 31] + hv[v & 31];
          if (cost < bc) {
            bc = cost;
            bi = i;
            if (cost == 0) break;
          }
        }
        w = R[bi];
        R[bi] = R.back();
        R.pop_back();
        qv[(w >> 16) & 15]++;
        hv[(w >> 16) & 31]++;
      }
      qj[w & 15]++;
      hj[w & 31]++;
      out[(size_t)64 * k + l] = w;
    }
  }
}

// Sliced-ELL layout of the singles of t with rows taken in `order`
// (64 rows per slice); results in host vectors.
static void ell_layout(Context &c, const SpinTables &t, const std::vector<uint64_t> &off,
                       const std::vector<uint64_t> &ent, const std::vector<uint32_t> &order,
                       std::vector<uint32_t> &perm, std::vector<uint32_t> &sl_len,
                       std::vector<uint64_t> &sl_off, std::vector<uint32_t> &ell) {
  const int n = t.n;
  const int nsl = (n + 63) / 64;
  perm.assign((size_t)nsl * 64, 0xffffffffu);
  sl_len.assign(nsl, 0);
  sl_off.assign(nsl, 0);
  uint64_t total = 0;
  for (int s = 0; s < nsl; s++) {
    uint32_t mx = 0;
    for (int l = 0; l < 64 && s * 64 + l < n; l++) {
      const uint32_t r = order[s * 64 + l];
      perm[(size_t)s * 64 + l] = r;
      mx = std::max<uint32_t>(mx, (uint32_t)(off[r + 1] - off[r]));
    }
    sl_len[s] = mx;
    sl_off[s] = total;
    total += (uint64_t)mx * 64;
  }
  const uint32_t pad = (uint32_t)c.K << 16;
  // 4 x 64 padding entries past the end: the mixed kernel prefetches the
  // next 4 positions of a slice unconditionally (past a slice's end it
  // reads the next slice's entries or these, and never uses them)
  ell.assign(total + 4 * 64 + 1, pad);
  std::vector<uint32_t> rows[64];
  for (int s = 0; s < nsl; s++) {
    for (int l = 0; l < 64; l++) {
      rows[l].clear();
      const uint32_t r = perm[(size_t)s * 64 + l];
      if (r == 0xffffffffu) continue;
      for (uint64_t k = 0; k < off[r + 1] - off[r]; k++) {
        const uint64_t e = ent[off[r] + k];
        const uint32_t ket = (uint32_t)(e & 0xffffffffull);
        const uint32_t pq = (uint32_t)((e >> 32) & 0x7fffffffull);
        rows[l].push_back(ket | (pq << 16) | ((e >> 63) ? 0x80000000u : 0u));
      }
    }
    schedule_slice(rows, sl_len[s], (uint32_t)c.K, (uint32_t)n, ell.data() + sl_off[s]);
  }
}

// Packed entries of the staged mixed kernel (SpinTables::ell2): bits 4..18
// hold K + ket, the LDS element of the ket's staged value (the staged ERI
// row occupies elements 0..K-1, the ket row K..K+nb-1, a zero at K+nb),
// so (w & 0x7fff0) is its byte offset for 16-byte elements; bits 19..30
// hold pq, (w >> 15) & 0xfff0 its byte offset; bit 31 the sign.  Padding
// entries read the zero ket element and ERI element 0, one address shared
// by every padding lane (an LDS broadcast).
//
// Dealing a slice's entries to positions.  Measured on gfx950 with the
// SQ_LDS_BANK_CONFLICT counter over per-lane index patterns (lanes in the
// same coset of a 2-dimensional subspace of lane-index bits sharing a
// residue, for every pair of bits): a 16-byte LDS read of a wavefront takes
// 4 cycles when conflict-free, lane l being served in lane group
// 2 * bit5(l) + (bit4 ^ bit3 ^ bit2)(l) (16 lanes per group), and two
// distinct addresses of one group conflict when their 16-byte element
// indices agree mod 16 (64 banks of 4 bytes); the cycles of one read are the
// sum over the 4 groups of the largest number of distinct addresses sharing
// a residue.  The groups are
// independent: for each, every position takes one entry per lane (or, for a
// lane with fewer entries than positions left, the shared padding address),
// and a depth-first search (the lane with the fewest options first, entries
// with the most heavily loaded residues first, a node budget) looks for a
// position whose ket residues and ERI residues each repeat at most (cx, cw)
// times in the group, trying (1,1), then (1,2) and (2,1), and so on; an
// unbounded greedy choice ends it.  On the norb-36 workload this takes the
// modelled cost of one read from 8.9 cycles (dealing that ignores the lane
// grouping) to 5.2 (4 without conflicts).  4- and 8-byte reads follow
// another rule (schedule_slice_packed), dealt by the same search.
namespace qsched {
// NL lanes per group, NR residues (a power of two, <= 32; NL == NR here)
template <int NR>
struct Lane {
  std::vector<uint32_t> ent;
  uint16_t cnt[NR * NR];  // entries per (x, w) residue pair
  uint32_t xw[NR];        // bit w of xw[x]: some entry has residues (x, w)
};
template <int NR>
static inline int rx(uint32_t w) { return (w >> 4) & (NR - 1); }
template <int NR>
static inline int rw(uint32_t w) { return (w >> 19) & (NR - 1); }
template <int NL, int NR>
struct Search {
  Lane<NR> *L;
  int remx[NR], remw[NR];
  uint8_t ux[NR], uw[NR];
  int cx, cw;
  int padx, padw;
  bool padused;
  long nodes, limit;
  int pick[NL];  // pair index x*NR+w, or -1 idle
  bool need[NL], done[NL];
  bool fits(int x, int w) const { return ux[x] < cx && uw[w] < cw; }
  int options(int l, uint32_t wfree) const {
    int n = 0;
    for (int x = 0; x < NR; x++)
      if (ux[x] < cx) n += __builtin_popcount(L[l].xw[x] & wfree);
    return n;
  }
  uint32_t wfree_mask() const {
    uint32_t m = 0;
    for (int w = 0; w < NR; w++) if (uw[w] < cw) m |= 1u << w;
    return m;
  }
  bool dfs(int depth) {
    if (depth == NL) return true;
    if (++nodes > limit) return false;
    int bl = -1, bn = 1 << 30;
    const uint32_t wf = wfree_mask();
    for (int l = 0; l < NL; l++) {
      if (done[l]) continue;
      int n = options(l, wf) + (need[l] ? 0 : 1);
      if (n < bn) { bn = n; bl = l; }
    }
    if (bn == 0) return false;
    const int l = bl;
    done[l] = true;
    std::array<std::pair<int, int>, NR * NR> cand;
    int nc = 0;
    for (int x = 0; x < NR; x++) {
      if (ux[x] >= cx) continue;
      uint32_t m = L[l].xw[x] & wf;
      while (m) {
        const int w = __builtin_ctz(m);
        m &= m - 1;
        cand[nc++] = {-(remx[x] + remw[w]), x * NR + w};
      }
    }
    std::sort(cand.begin(), cand.begin() + nc);
    for (int i = 0; i < nc; i++) {
      const int p = cand[i].second, x = p / NR, w = p % NR;
      ux[x]++; uw[w]++;
      pick[l] = p;
      if (dfs(depth + 1)) return true;
      ux[x]--; uw[w]--;
      if (nodes > limit) break;
    }
    if (!need[l]) {
      const bool was = padused;
      if (was || fits(padx, padw)) {
        if (!was) { ux[padx]++; uw[padw]++; padused = true; }
        pick[l] = -1;
        if (dfs(depth + 1)) return true;
        if (!was) { ux[padx]--; uw[padw]--; padused = false; }
      }
    }
    done[l] = false;
    return false;
  }
};
// rows[lanes[i]]: the NL lanes' entries (consumed); out[64*k + lanes[i]]
template <int NL, int NR>
static void deal_group(std::vector<uint32_t> *rows, const int *lanes, uint32_t len, uint32_t pad, uint32_t *out,
                       long limit) {
  std::vector<Lane<NR>> L(NL);
  Search<NL, NR> s;
  s.L = L.data();
  std::fill(s.remx, s.remx + NR, 0);
  std::fill(s.remw, s.remw + NR, 0);
  for (int l = 0; l < NL; l++) {
    L[l].ent = rows[lanes[l]];
    std::fill(L[l].cnt, L[l].cnt + NR * NR, 0);
    std::fill(L[l].xw, L[l].xw + NR, 0);
    for (uint32_t w : L[l].ent) {
      L[l].cnt[rx<NR>(w) * NR + rw<NR>(w)]++;
      L[l].xw[rx<NR>(w)] |= 1u << rw<NR>(w);
      s.remx[rx<NR>(w)]++;
      s.remw[rw<NR>(w)]++;
    }
  }
  s.padx = rx<NR>(pad);
  s.padw = rw<NR>(pad);
  static const int caps[][2] = {{1, 1}, {1, 2}, {2, 1}, {2, 2}, {2, 3}, {3, 2}, {3, 3}};
  const int ncaps = sizeof(caps) / sizeof(caps[0]);
  for (uint32_t k = 0; k < len; k++) {
    const uint32_t left = len - k;
    for (int l = 0; l < NL; l++) s.need[l] = L[l].ent.size() >= left;
    bool ok = false;
    for (int c = 0; c < ncaps && !ok; c++) {
      s.cx = caps[c][0];
      s.cw = caps[c][1];
      std::fill(s.ux, s.ux + NR, 0);
      std::fill(s.uw, s.uw + NR, 0);
      std::fill(s.done, s.done + NL, false);
      s.padused = false;
      s.nodes = 0;
      s.limit = limit;
      ok = s.dfs(0);
    }
    if (!ok) {
      // unbounded: lane by lane, the pair minimizing the larger use count
      std::fill(s.ux, s.ux + NR, 0);
      std::fill(s.uw, s.uw + NR, 0);
      bool pu = false;
      for (int l = 0; l < NL; l++) {
        if (L[l].ent.empty()) {
          s.pick[l] = -1;
          if (!pu) { s.ux[s.padx]++; s.uw[s.padw]++; pu = true; }
          continue;
        }
        long best = 1L << 60;
        int bp = 0;
        for (int p = 0; p < NR * NR; p++) {
          if (!L[l].cnt[p]) continue;
          const int x = p / NR, w = p % NR;
          const long sc = (long)std::max(s.ux[x], s.uw[w]) * 1000000 + (long)(s.ux[x] + s.uw[w]) * 1000 -
                          (s.remx[x] + s.remw[w]);
          if (sc < best) { best = sc; bp = p; }
        }
        s.pick[l] = bp;
        s.ux[bp / NR]++;
        s.uw[bp % NR]++;
      }
    }
    for (int l = 0; l < NL; l++) {
      uint32_t w = pad;
      if (s.pick[l] >= 0 && !L[l].ent.empty()) {
        const int p = s.pick[l];
        auto &E = L[l].ent;
        for (size_t i = 0; i < E.size(); i++)
          if (rx<NR>(E[i]) * NR + rw<NR>(E[i]) == p) { w = E[i]; E[i] = E.back(); E.pop_back(); break; }
        if (--L[l].cnt[p] == 0) L[l].xw[p / NR] &= ~(1u << (p % NR));
        s.remx[p / NR]--;
        s.remw[p % NR]--;
      }
      out[(size_t)64 * k + lanes[l]] = w;
    }
  }
}
}  // namespace qsched

// The packed entries of one slice (rows: the 64 lanes' entries, consumed),
// group by group: for 16-byte reads the 4 measured lane groups of 16 lanes
// with residues mod 16; for 4- and 8-byte reads the two half-waves with
// residues mod 32 (measured with the same probe: lanes of different
// half-waves never conflict, and two distinct addresses of one half-wave
// conflict when their element indices agree mod 32).
static void schedule_slice_packed(std::vector<uint32_t> *rows, uint32_t len, uint32_t pad, uint32_t *out,
                                  bool wide) {
  if (wide) {
    for (int c = 0; c < 4; c++) {
      int lanes[16], n = 0;
      for (int l = 0; l < 64; l++)
        if (((l >> 5) & 1) * 2 + (((l >> 4) ^ (l >> 3) ^ (l >> 2)) & 1) == c) lanes[n++] = l;
      qsched::deal_group<16, 16>(rows, lanes, len, pad, out, 200);
    }
  } else {
    for (int c = 0; c < 2; c++) {
      int lanes[32];
      for (int l = 0; l < 32; l++) lanes[l] = 32 * c + l;
      qsched::deal_group<32, 32>(rows, lanes, len, pad, out, 200);
    }
  }
}

static void download_singles(Context &c, const SpinTables &t, std::vector<uint64_t> &off,
                             std::vector<uint64_t> &ent) {
  off.resize(t.n + 1);
  ent.resize(t.ns);
  SBDX_HIP(hipMemcpyAsync(off.data(), t.s_off, sizeof(uint64_t) * (t.n + 1), hipMemcpyDeviceToHost, c.stream));
  SBDX_HIP(hipMemcpyAsync(ent.data(), t.s_ent, sizeof(uint64_t) * t.ns, hipMemcpyDeviceToHost, c.stream));
  SBDX_HIP(hipStreamSynchronize(c.stream));
}

void build_singles_ell(Context &c, SpinTables &t) {
  t.has_ell = false;
  if (t.n > 65536 || c.K >= 32768) return;
  const int n = t.n;
  std::vector<uint64_t> off, ent;
  download_singles(c, t, off, ent);
  // the singles grouped by pair index (counting sort, row order kept
  // within a pair) for the RDM's mixed term
  {
    const size_t K = (size_t)c.K;
    std::vector<uint64_t> poff(K + 1, 0);
    for (size_t k = 0; k < t.ns; k++) poff[((ent[k] >> 32) & 0x7fffffffull) + 1]++;
    for (size_t pq = 0; pq < K; pq++) poff[pq + 1] += poff[pq];
    std::vector<uint64_t> pent(t.ns);
    std::vector<uint64_t> fill(poff.begin(), poff.end() - 1);
    for (int r = 0; r < n; r++)
      for (uint64_t k = off[r]; k < off[r + 1]; k++) {
        const uint64_t e = ent[k];
        const uint64_t dst = fill[(e >> 32) & 0x7fffffffull]++;
        const uint32_t ket = (uint32_t)(e & 0xffffffffull), neg = (uint32_t)(e >> 63);
        pent[dst] = (uint64_t)r | ((uint64_t)ket << 32) | ((uint64_t)neg << 63);
      }
    t.pm_off = (uint64_t *)dalloc(c, sizeof(uint64_t) * (K + 1));
    t.pm_ent = (uint64_t *)dalloc(c, sizeof(uint64_t) * std::max<size_t>(t.ns, 1));
    SBDX_HIP(hipMemcpyAsync(t.pm_off, poff.data(), sizeof(uint64_t) * (K + 1), hipMemcpyHostToDevice, c.stream));
    SBDX_HIP(hipMemcpyAsync(t.pm_ent, pent.data(), sizeof(uint64_t) * t.ns, hipMemcpyHostToDevice, c.stream));
    // the same grouping as a sliced ELL over the pair bins (n <= 32767):
    // bins sorted by entry count (descending) into slices of 64 lanes,
    // entry k of lane l of slice s at pb_off[s] + 64 k + l, packed as
    // bra | ket << 15 | (sign<0) << 31; padding entries bra = ket = n (a
    // staged zero); pb_perm[64 s + l] = the lane's bin (0xffffffff past K)
    t.pb_nslice = 0;
    if (n < 32768) {
      std::vector<uint32_t> bins(K);
      for (size_t i = 0; i < K; i++) bins[i] = (uint32_t)i;
      std::stable_sort(bins.begin(), bins.end(), [&](uint32_t a, uint32_t b) {
        return (poff[a + 1] - poff[a]) > (poff[b + 1] - poff[b]);
      });
      const size_t nsl = (K + 63) / 64;
      std::vector<uint32_t> perm(nsl * 64, 0xffffffffu), len(nsl);
      std::vector<uint64_t> soff(nsl);
      uint64_t tot = 0;
      for (size_t sl = 0; sl < nsl; sl++) {
        uint64_t mx = 0;
        for (size_t l = 0; l < 64; l++) {
          const size_t slot = sl * 64 + l;
          if (slot < K) {
            perm[slot] = bins[slot];
            mx = std::max<uint64_t>(mx, poff[bins[slot] + 1] - poff[bins[slot]]);
          }
        }
        soff[sl] = tot;
        len[sl] = (uint32_t)mx;
        tot += mx * 64;
      }
      const uint32_t pad = (uint32_t)n | ((uint32_t)n << 15);
      std::vector<uint32_t> pb(std::max<uint64_t>(tot, 1), pad);
      for (size_t sl = 0; sl < nsl; sl++)
        for (size_t l = 0; l < 64; l++) {
          const uint32_t bin = perm[sl * 64 + l];
          if (bin == 0xffffffffu) continue;
          for (uint64_t k = poff[bin]; k < poff[bin + 1]; k++) {
            const uint64_t pe = pent[k];
            const uint32_t bra = (uint32_t)(pe & 0xffffffffull), ket = (uint32_t)((pe >> 32) & 0x7fffffffull);
            pb[soff[sl] + 64 * (k - poff[bin]) + l] = bra | (ket << 15) | ((uint32_t)(pe >> 63) << 31);
          }
        }
      t.pb = (uint32_t *)dalloc(c, sizeof(uint32_t) * pb.size());
      t.pb_off = (uint64_t *)dalloc(c, sizeof(uint64_t) * nsl);
      t.pb_len = (uint32_t *)dalloc(c, sizeof(uint32_t) * nsl);
      t.pb_perm = (uint32_t *)dalloc(c, sizeof(uint32_t) * perm.size());
      SBDX_HIP(hipMemcpyAsync(t.pb, pb.data(), sizeof(uint32_t) * pb.size(), hipMemcpyHostToDevice, c.stream));
      SBDX_HIP(hipMemcpyAsync(t.pb_off, soff.data(), sizeof(uint64_t) * nsl, hipMemcpyHostToDevice, c.stream));
      SBDX_HIP(hipMemcpyAsync(t.pb_len, len.data(), sizeof(uint32_t) * nsl, hipMemcpyHostToDevice, c.stream));
      SBDX_HIP(hipMemcpyAsync(t.pb_perm, perm.data(), sizeof(uint32_t) * perm.size(), hipMemcpyHostToDevice,
                              c.stream));
      t.pb_nslice = (int)nsl;
    }
    SBDX_HIP(hipStreamSynchronize(c.stream));
  }
  std::vector<uint32_t> order(n);
  for (int i = 0; i < n; i++) order[i] = (uint32_t)i;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return (off[a + 1] - off[a]) > (off[b + 1] - off[b]);
  });
  std::vector<uint32_t> perm, sl_len, ell;
  std::vector<uint64_t> sl_off;
  ell_layout(c, t, off, ent, order, perm, sl_len, sl_off, ell);
  const int nsl = (int)sl_len.size();
  t.nslice = nsl;
  t.ell = (uint32_t *)dalloc(c, sizeof(uint32_t) * ell.size());
  t.perm = (uint32_t *)dalloc(c, sizeof(uint32_t) * perm.size());
  t.sl_off = (uint64_t *)dalloc(c, sizeof(uint64_t) * nsl);
  t.sl_len = (uint32_t *)dalloc(c, sizeof(uint32_t) * nsl);
  SBDX_HIP(hipMemcpyAsync(t.ell, ell.data(), sizeof(uint32_t) * ell.size(), hipMemcpyHostToDevice, c.stream));
  SBDX_HIP(hipMemcpyAsync(t.perm, perm.data(), sizeof(uint32_t) * perm.size(), hipMemcpyHostToDevice, c.stream));
  SBDX_HIP(hipMemcpyAsync(t.sl_off, sl_off.data(), sizeof(uint64_t) * nsl, hipMemcpyHostToDevice, c.stream));
  SBDX_HIP(hipMemcpyAsync(t.sl_len, sl_len.data(), sizeof(uint32_t) * nsl, hipMemcpyHostToDevice, c.stream));
  SBDX_HIP(hipStreamSynchronize(c.stream));
  t.has_ell = true;
  // the packed form for k_mixed_ell (same rows, slices and lengths)
  t.has_ell2 = false;
  t.has_ell3 = false;
  if (c.K > 4096 || (size_t)c.K + (size_t)n > 32767) return;
  const uint32_t pad2 = ((uint32_t)(c.K + n)) << 4;
  // the reads the kernel will make: 16-byte values when four FP32 pairs fit
  // in LDS (the autotuner's choice whenever they do), else 4- or 8-byte ones
  const bool wide = (size_t)4 * ((size_t)c.K + n + 1) * sizeof(float) <= c.lds_limit;
  // groups of 4 positions: slice lengths rounded up to a multiple of 4, each
  // group stored lane-major (lane l's 4 words contiguous, one 16-byte load),
  // plus 2 groups of padding past the end for the kernel's prefetch
  std::vector<uint32_t> sl_len2(nsl);
  std::vector<uint64_t> sl_off2(nsl);
  uint64_t total2 = 0;
  for (int sl = 0; sl < nsl; sl++) {
    sl_len2[sl] = (sl_len[sl] + 3) & ~3u;
    sl_off2[sl] = total2;
    total2 += (uint64_t)sl_len2[sl] * 64;
  }
  // deal every slice's entries (word_of: the packed word of one singles
  // entry, pad the padding word) into the group layout; slices are
  // independent and dealt on up to 16 host threads
  auto deal_all = [&](auto word_of, uint32_t pad, std::vector<uint32_t> &dst) {
    dst.assign(total2 + 2 * 256, pad);
    const int nthr = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::atomic<int> next{0};
    auto worker = [&]() {
      std::vector<uint32_t> rows[64];
      std::vector<uint32_t> dealt;
      for (int sl = next++; sl < nsl; sl = next++) {
        for (int l = 0; l < 64; l++) {
          rows[l].clear();
          const uint32_t r = perm[(size_t)sl * 64 + l];
          if (r == 0xffffffffu) continue;
          for (uint64_t k = off[r]; k < off[r + 1]; k++) rows[l].push_back(word_of(ent[k]));
        }
        const uint32_t len2 = sl_len2[sl];
        dealt.assign((size_t)len2 * 64, pad);
        schedule_slice_packed(rows, sl_len[sl], pad, dealt.data(), wide);
        uint32_t *o = dst.data() + sl_off2[sl];
        for (uint32_t k = 0; k < len2; k++)
          for (int l = 0; l < 64; l++) o[256 * (k / 4) + 4 * l + (k % 4)] = dealt[(size_t)64 * k + l];
      }
    };
    std::vector<std::thread> pool;
    for (int i = 1; i < nthr; i++) pool.emplace_back(worker);
    worker();
    for (auto &th : pool) th.join();
  };
  std::vector<uint32_t> ell2;
  deal_all(
      [&](uint64_t e) {
        const uint32_t ket = (uint32_t)(e & 0xffffffffull);
        const uint32_t pq = (uint32_t)((e >> 32) & 0x7fffffffull);
        return ((uint32_t)(c.K + ket) << 4) | (pq << 19) | ((e >> 63) ? 0x80000000u : 0u);
      },
      pad2, ell2);
  // the signed-triangle form: dealt in the ell2 word layout with the ket
  // element at 2048 + ket (the kernel's ket rows start at byte 32768 = 2048
  // 16-byte elements, so residues match the LDS addresses) and the signed
  // triangle index in the pq field, then repacked
  if (c.eri_rs_sym && c.Kt <= 1024 && n < 8192 && wide) {
    const uint32_t Kt = (uint32_t)c.Kt;
    const int L = c.L;
    std::vector<uint32_t> ell3;
    deal_all(
        [&](uint64_t e) {
          const uint32_t ket = (uint32_t)(e & 0xffffffffull);
          const uint32_t pq = (uint32_t)((e >> 32) & 0x7fffffffull);
          const uint32_t a = pq / L, b = pq % L, hi = a > b ? a : b, lo = a > b ? b : a;
          const uint32_t widx = hi * (hi + 1) / 2 + lo + ((e >> 63) ? Kt : 0u);
          return ((2048u + ket) << 4) | (widx << 19);
        },
        (2048u + (uint32_t)n) << 4, ell3);
    for (uint32_t &w : ell3) w = ((((w >> 4) & 0x7fffu) - 2048u) << 4) | (((w >> 19) & 0xfffu) << 21);
    t.ell3 = (uint32_t *)dalloc(c, sizeof(uint32_t) * ell3.size());
    SBDX_HIP(hipMemcpyAsync(t.ell3, ell3.data(), sizeof(uint32_t) * ell3.size(), hipMemcpyHostToDevice,
                            c.stream));
    SBDX_HIP(hipStreamSynchronize(c.stream));
    t.has_ell3 = true;
  }
  t.ell2 = (uint32_t *)dalloc(c, sizeof(uint32_t) * ell2.size());
  t.sl_off2 = (uint64_t *)dalloc(c, sizeof(uint64_t) * nsl);
  t.sl_len2 = (uint32_t *)dalloc(c, sizeof(uint32_t) * nsl);
  SBDX_HIP(hipMemcpyAsync(t.ell2, ell2.data(), sizeof(uint32_t) * ell2.size(), hipMemcpyHostToDevice, c.stream));
  SBDX_HIP(hipMemcpyAsync(t.sl_off2, sl_off2.data(), sizeof(uint64_t) * nsl, hipMemcpyHostToDevice, c.stream));
  SBDX_HIP(hipMemcpyAsync(t.sl_len2, sl_len2.data(), sizeof(uint32_t) * nsl, hipMemcpyHostToDevice, c.stream));
  SBDX_HIP(hipStreamSynchronize(c.stream));
  t.has_ell2 = true;
}

// Ket-range parts of the sliced-ELL singles (SpinTables::pell): part p holds
// the entries with ket in [p*h, (p+1)*h), h = ceil(n/nparts), ket stored as
// ket - p*h, with the slices and row order of t.perm.
void build_singles_ell_parts(Context &c, SpinTables &t, int nparts) {
  if (!t.has_ell || nparts < 2 || t.nparts == nparts) return;
  const int n = t.n;
  const int h = (n + nparts - 1) / nparts;
  std::vector<uint64_t> off(n + 1), ent(t.ns);
  std::vector<uint32_t> perm((size_t)t.nslice * 64);
  SBDX_HIP(hipMemcpyAsync(off.data(), t.s_off, sizeof(uint64_t) * (n + 1), hipMemcpyDeviceToHost, c.stream));
  SBDX_HIP(hipMemcpyAsync(ent.data(), t.s_ent, sizeof(uint64_t) * t.ns, hipMemcpyDeviceToHost, c.stream));
  SBDX_HIP(hipMemcpyAsync(perm.data(), t.perm, sizeof(uint32_t) * perm.size(), hipMemcpyDeviceToHost, c.stream));
  SBDX_HIP(hipStreamSynchronize(c.stream));
  const uint32_t pad = (uint32_t)c.K << 16;
  t.pell.assign(nparts, nullptr);
  t.psl_off.assign(nparts, nullptr);
  t.psl_len.assign(nparts, nullptr);
  for (int p = 0; p < nparts; p++) {
    const uint64_t lo = (uint64_t)p * h, hi = std::min<uint64_t>((uint64_t)(p + 1) * h, (uint64_t)n);
    std::vector<uint32_t> sl_len(t.nslice);
    std::vector<uint64_t> sl_off(t.nslice);
    uint64_t total = 0;
    for (int s = 0; s < t.nslice; s++) {
      uint32_t mx = 0;
      for (int l = 0; l < 64; l++) {
        const uint32_t r = perm[(size_t)s * 64 + l];
        if (r == 0xffffffffu) continue;
        uint32_t cnt = 0;
        for (uint64_t e = off[r]; e < off[r + 1]; e++) {
          const uint64_t k = ent[e] & 0xffffffffull;
          cnt += (k >= lo && k < hi);
        }
        mx = std::max(mx, cnt);
      }
      sl_len[s] = mx;
      sl_off[s] = total;
      total += (uint64_t)mx * 64;
    }
    std::vector<uint32_t> ell(total + 1, pad);
    std::vector<uint32_t> rows[64];
    for (int s = 0; s < t.nslice; s++) {
      for (int l = 0; l < 64; l++) {
        rows[l].clear();
        const uint32_t r = perm[(size_t)s * 64 + l];
        if (r == 0xffffffffu) continue;
        for (uint64_t e = off[r]; e < off[r + 1]; e++) {
          const uint64_t ket = ent[e] & 0xffffffffull;
          if (ket < lo || ket >= hi) continue;
          const uint32_t pq = (uint32_t)((ent[e] >> 32) & 0x7fffffffull);
          rows[l].push_back((uint32_t)(ket - lo) | (pq << 16) | ((ent[e] >> 63) ? 0x80000000u : 0u));
        }
      }
      schedule_slice(rows, sl_len[s], (uint32_t)c.K, (uint32_t)(hi - lo), ell.data() + sl_off[s]);
    }
    t.pell[p] = (uint32_t *)dalloc(c, sizeof(uint32_t) * ell.size());
    t.psl_off[p] = (uint64_t *)dalloc(c, sizeof(uint64_t) * t.nslice);
    t.psl_len[p] = (uint32_t *)dalloc(c, sizeof(uint32_t) * t.nslice);
    SBDX_HIP(hipMemcpyAsync(t.pell[p], ell.data(), sizeof(uint32_t) * ell.size(), hipMemcpyHostToDevice, c.stream));
    SBDX_HIP(hipMemcpyAsync(t.psl_off[p], sl_off.data(), sizeof(uint64_t) * t.nslice, hipMemcpyHostToDevice, c.stream));
    SBDX_HIP(hipMemcpyAsync(t.psl_len[p], sl_len.data(), sizeof(uint32_t) * t.nslice, hipMemcpyHostToDevice, c.stream));
    SBDX_HIP(hipStreamSynchronize(c.stream));
  }
  t.nparts = nparts;
  t.part_h = h;
}

void compute_diag(Context &c) {
  const unsigned nchunk = cdiv(c.nb, 256);
  ProfScope ps(c, "diag");
  hipLaunchKernelGGL(k_diag, dim3((unsigned)(c.na * nchunk)), dim3(256), 0, c.stream, (int)c.nb,
                     (int)nchunk, c.L, c.I0, (c.B.alias && c.na == c.nb) ? 1 : 0, c.B.str, c.A.e1, c.B.e1,
                     c.A.jsum, c.hii);
  check_launch("k_diag");
}

void sign_convert(Context &c, const double *in, double *out) {
  const unsigned nchunk = cdiv(c.nb, 256);
  ProfScope ps(c, "sign");
  hipLaunchKernelGGL(k_sign<false>, dim3((unsigned)(c.na * nchunk)), dim3(256), 0, c.stream,
                     (int)c.nb, (int)nchunk, c.A.str, c.B.str, in, out);
  check_launch("k_sign");
}

void sign_convert_noprof(Context &c, const double *in, double *out) {
  const unsigned nchunk = cdiv(c.nb, 256);
  hipLaunchKernelGGL(k_sign<false>, dim3((unsigned)(c.na * nchunk)), dim3(256), 0, c.stream,
                     (int)c.nb, (int)nchunk, c.A.str, c.B.str, in, out);
  check_launch("k_sign");
}

void sign_convert_add(Context &c, const double *in, double *out) {
  const unsigned nchunk = cdiv(c.nb, 256);
  ProfScope ps(c, "sign");
  hipLaunchKernelGGL(k_sign<true>, dim3((unsigned)(c.na * nchunk)), dim3(256), 0, c.stream,
                     (int)c.nb, (int)nchunk, c.A.str, c.B.str, in, out);
  check_launch("k_sign");
}

}  // namespace sbdx
