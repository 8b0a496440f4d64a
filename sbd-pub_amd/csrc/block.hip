// Block Davidson on the device: the lowest nroots eigenpairs with up to
// `block` correction vectors per iteration pushed through one block sigma.
// No counterpart in the reference, whose Davidson finds one root
// (chemistry/tpb/davidson.h:667-930); the per-root steps (Ritz pair,
// residual, diagonal preconditioner (E - d)^-1 with the 1e-12 guard of
// davidson.h:861-868) are the reference's, applied to a block.
//
// Storage.  The basis C and its products HC are columns of two banks: an
// FP64 bank and an FP32 bank.  In FP64 mode every column is FP64.  In mixed
// mode the first nroots columns (the roots after a thick restart, with exact
// FP64 products) are FP64 and every correction column is FP32 in storage,
// as its product is in arithmetic: the correction's FP32 product is stored
// as computed, so FP32 storage adds no error to that pair.  Every dot
// product and combination reads the stored values into FP64 arithmetic.
// Since rounding a column to FP32 after orthogonalization leaves it
// orthogonal only to ~1e-7, the Rayleigh-Ritz step solves the generalized
// problem H y = theta S y with the overlap S = C^T C of the stored columns
// (Cholesky of S, then the symmetric eigensolver, on the host: at most
// 64 x 64).  At c58 (400M determinants) this halves the basis: 4 FP64 roots
// plus 44 FP32 columns take 167 GB where 48 FP64 columns would take 307 GB.
//
// Per iteration: Rayleigh-Ritz; one pass forms the lowest `block` Ritz
// pairs' residuals R_j = HC u_j - theta_j C u_j from the basis (no Ritz
// vectors are stored) and preconditions them; roots whose residual is below
// eps are locked (no correction); classical Gram-Schmidt twice against the
// basis, modified Gram-Schmidt inside the block (directions whose norm
// falls below 1e-10 of their norm before projection are dropped); one block
// product; the new columns of H = HC^T C (entry i <= j from the older
// column's product HC_i . C_j) and of S.  Thick restart when the next block
// would not fit: the basis becomes the lowest `keep` Ritz vectors and their
// products, combined in place; in mixed mode, once max|r| is below 100 eps
// or near the FP32 products' error level, the roots get exact FP64
// products at every restart.  Convergence: |R_j| < eps
// for every root on a basis whose columns and products are all FP64 (mixed
// mode: a confirming restart keeps the roots alone, with exact products).
//
// Every reduction is a fixed tree (fixed grid, fixed shuffle pattern, fixed
// order over blocks), so a solve is reproducible run to run.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <string>

#include "common.h"

namespace sbdx {

namespace {

constexpr int kBT = 256;      // threads per block of the vector kernels
constexpr int kBG = 1024;     // blocks: a fixed grid, so reductions are fixed trees
constexpr int kOutMax = 32;   // columns written by one combination pass
constexpr int kResMax = 16;   // residuals of one pass (the largest block)
constexpr int kCoarseIt = 25;  // iteration cap of the coarse stage
// error level of an FP32 product relative to |E| (1000 x 2^-24)
constexpr double kFp32Level = 1000.0 * 5.9604644775390625e-8;

// Basis columns: the first nd in FP64 storage, then nf in FP32 storage.
struct Cols {
  const double *d[kNbMax];
  const float *f[kNbMax];
  int nd, nf;
};
struct OutCols {
  double *d[kOutMax];
  float *f[kOutMax];
  int nd, nf;
};
struct OutD {  // FP64 vectors; a null entry is skipped
  double *p[kResMax];
};

__device__ __forceinline__ double wave_sum(double v) {
  // xor butterfly: every lane ends with the same, fixed-order sum
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Wave w's sum of v to sh[w*stride + idx]; after a barrier the block sum is
// ((sh[idx] + sh[stride+idx]) + sh[2 stride+idx]) + sh[3 stride+idx].
// Called by every thread of the block (4 waves).
__device__ __forceinline__ void block_put(double v, int idx, double *sh, int stride) {
  const double w = wave_sum(v);
  if ((threadIdx.x & 63) == 0) sh[(threadIdx.x >> 6) * stride + idx] = w;
}

template <typename TA, typename TB>
__device__ __forceinline__ void gram_tile(const TA *const *pa, int ma, const TB *const *pb, int mb, size_t n,
                                          double (&s)[8][8]) {
  const TA *a[8];
  const TB *b[8];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    a[q] = pa[q < ma ? q : 0];
    b[q] = pb[q < mb ? q : 0];
  }
#pragma unroll
  for (int x = 0; x < 8; x++)
#pragma unroll
    for (int y = 0; y < 8; y++) s[x][y] = 0.0;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double av[8], bv[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      av[q] = (double)a[q][i];
      bv[q] = (double)b[q][i];
    }
#pragma unroll
    for (int x = 0; x < 8; x++)
#pragma unroll
      for (int y = 0; y < 8; y++) s[x][y] += av[x] * bv[y];
  }
}

// part[blk * (na*nb) + a*nb + b] = this block's share of A_a . B_b, for the
// na = A.nd + A.nf and nb = B.nd + B.nf columns (tiles of 8 x 8 that do not
// straddle the FP64/FP32 boundary of either list).
__global__ void __launch_bounds__(kBT) k_bgram(Cols A, Cols B, size_t n, double *__restrict__ part) {
  __shared__ double sh[4 * 64];
  const int na = A.nd + A.nf, nb = B.nd + B.nf, tot = na * nb;
  for (int sa = 0; sa < 2; sa++) {
    const int nsa = sa ? A.nf : A.nd, offa = sa ? A.nd : 0;
    for (int ga = 0; ga < nsa; ga += 8) {
      const int ma = min(8, nsa - ga);
      for (int sb = 0; sb < 2; sb++) {
        const int nsb = sb ? B.nf : B.nd, offb = sb ? B.nd : 0;
        for (int gb = 0; gb < nsb; gb += 8) {
          const int mb = min(8, nsb - gb);
          double s[8][8];
          if (!sa && !sb) gram_tile<double, double>(A.d + ga, ma, B.d + gb, mb, n, s);
          else if (!sa) gram_tile<double, float>(A.d + ga, ma, B.f + gb, mb, n, s);
          else if (!sb) gram_tile<float, double>(A.f + ga, ma, B.d + gb, mb, n, s);
          else gram_tile<float, float>(A.f + ga, ma, B.f + gb, mb, n, s);
#pragma unroll
          for (int x = 0; x < 8; x++)
#pragma unroll
            for (int y = 0; y < 8; y++) block_put(s[x][y], x * 8 + y, sh, 64);
          __syncthreads();
          const int t = threadIdx.x;
          if (t < 64 && (t >> 3) < ma && (t & 7) < mb) {
            const double v = ((sh[t] + sh[64 + t]) + sh[128 + t]) + sh[192 + t];
            part[(size_t)blockIdx.x * tot + (size_t)(offa + ga + (t >> 3)) * nb + offb + gb + (t & 7)] = v;
          }
          __syncthreads();
        }
      }
    }
  }
}

// out[t] = sum over blocks b (in order) of part[b*tot + t]
__global__ void k_bsum(int tot, int nblk, const double *__restrict__ part, double *__restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= tot) return;
  double s = 0.0;
  for (int b = 0; b < nblk; b++) s += part[(size_t)b * tot + t];
  out[t] = s;
}

// out_j = sum_k U[k + ld*j] in_k for j < nout <= kOutMax.  In place is
// allowed (out_j the same storage as in_j): each thread reads every input
// element before it writes its outputs.
__global__ void __launch_bounds__(kBT) k_bcomb(Cols in, const double *__restrict__ U, int ld, OutCols out,
                                               size_t n) {
  __shared__ double u[kOutMax * kNbMax];
  const int nin = in.nd + in.nf, nout = out.nd + out.nf;
  for (int t = threadIdx.x; t < nin * nout; t += blockDim.x) {
    const int k = t % nin, j = t / nin;
    u[j * kNbMax + k] = U[k + (size_t)ld * j];
  }
  __syncthreads();
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double acc[kOutMax];
#pragma unroll
    for (int j = 0; j < kOutMax; j++) acc[j] = 0.0;
    for (int s = 0; s < 2; s++) {
      const int ns = s ? in.nf : in.nd, off = s ? in.nd : 0;
      for (int k0 = 0; k0 < ns; k0 += 8) {  // loads of 8 columns issue together
        const int m = min(8, ns - k0);
        double v[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
          const int k = q < m ? k0 + q : k0;
          v[q] = s ? (double)in.f[k][i] : in.d[k][i];
        }
#pragma unroll
        for (int q = 0; q < 8; q++)
          if (q < m) {
            const double *uk = u + off + k0 + q;
#pragma unroll
            for (int j = 0; j < kOutMax; j++)
              if (j < nout) acc[j] += uk[j * kNbMax] * v[q];
          }
      }
    }
#pragma unroll
    for (int j = 0; j < kOutMax; j++)
      if (j < nout) {
        if (j < out.nd) out.d[j][i] = acc[j];
        else out.f[j - out.nd][i] = (float)acc[j];
      }
  }
}

// Residuals of the Ritz pairs j < nr <= kResMax:
//   w_j = sum_k U[k + ld*j] C_k,  hw_j = sum_k U[k + ld*j] HC_k,
//   r_j = hw_j - E_j w_j,
// written to out.p[j] (when not null) as r_j, or with precond as the
// diagonal step r_j / (E_j - d) (E_j - d - reg when |E_j - d| <= reg);
// part[blk*nr + j] = this block's share of |r_j|^2.
__global__ void __launch_bounds__(kBT) k_bresid(Cols C, Cols HC, const double *__restrict__ U, int ld,
                                                const double *__restrict__ E, int nr,
                                                const double *__restrict__ hii, double reg, int precond,
                                                OutD out, size_t n, double *__restrict__ part) {
  __shared__ double u[kResMax * kNbMax];
  __shared__ double e[kResMax];
  __shared__ double sh[4 * kResMax];
  const int nin = C.nd + C.nf;
  for (int t = threadIdx.x; t < nin * nr; t += blockDim.x) {
    const int k = t % nin, j = t / nin;
    u[j * kNbMax + k] = U[k + (size_t)ld * j];
  }
  if (threadIdx.x < nr) e[threadIdx.x] = E[threadIdx.x];
  __syncthreads();
  double ss[kResMax];
#pragma unroll
  for (int j = 0; j < kResMax; j++) ss[j] = 0.0;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double w[kResMax], hw[kResMax];
#pragma unroll
    for (int j = 0; j < kResMax; j++) w[j] = hw[j] = 0.0;
    for (int s = 0; s < 2; s++) {
      const int ns = s ? C.nf : C.nd, off = s ? C.nd : 0;
      for (int k0 = 0; k0 < ns; k0 += 4) {  // loads of 4 columns and their products issue together
        const int m = min(4, ns - k0);
        double cv[4], hv[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int k = q < m ? k0 + q : k0;
          cv[q] = s ? (double)C.f[k][i] : C.d[k][i];
          hv[q] = s ? (double)HC.f[k][i] : HC.d[k][i];
        }
#pragma unroll
        for (int q = 0; q < 4; q++)
          if (q < m) {
            const double *uk = u + off + k0 + q;
#pragma unroll
            for (int j = 0; j < kResMax; j++)
              if (j < nr) {
                w[j] += uk[j * kNbMax] * cv[q];
                hw[j] += uk[j * kNbMax] * hv[q];
              }
          }
      }
    }
    const double d = hii[i];
#pragma unroll
    for (int j = 0; j < kResMax; j++)
      if (j < nr) {
        const double r = hw[j] - e[j] * w[j];
        ss[j] += r * r;
        if (out.p[j]) {
          double v = r;
          if (precond) {
            const double den = e[j] - d;
            v = fabs(den) > reg ? r / den : r / (den - reg);
          }
          out.p[j][i] = v;
        }
      }
  }
#pragma unroll
  for (int j = 0; j < kResMax; j++) block_put(ss[j], j, sh, kResMax);
  __syncthreads();
  const int t = threadIdx.x;
  if (t < nr) part[(size_t)blockIdx.x * nr + t] = ((sh[t] + sh[kResMax + t]) + sh[2 * kResMax + t]) +
                                                  sh[3 * kResMax + t];
}

// t_j -= sum_k coef[k*nv + j] X_k for j < nv <= kResMax (one classical
// Gram-Schmidt pass with reduced coefficients)
__global__ void __launch_bounds__(kBT) k_bsub(Cols X, const double *__restrict__ coef, int nv, OutD t,
                                              size_t n) {
  __shared__ double cf[kNbMax * kResMax];
  const int nx = X.nd + X.nf;
  for (int q = threadIdx.x; q < nx * nv; q += blockDim.x) cf[q] = coef[q];
  __syncthreads();
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double acc[kResMax];
#pragma unroll
    for (int j = 0; j < kResMax; j++) acc[j] = j < nv ? t.p[j][i] : 0.0;
    for (int s = 0; s < 2; s++) {
      const int ns = s ? X.nf : X.nd, off = s ? X.nd : 0;
      for (int k0 = 0; k0 < ns; k0 += 8) {
        const int m = min(8, ns - k0);
        double v[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
          const int k = q < m ? k0 + q : k0;
          v[q] = s ? (double)X.f[k][i] : X.d[k][i];
        }
#pragma unroll
        for (int q = 0; q < 8; q++)
          if (q < m) {
            const double *ck = cf + (size_t)(off + k0 + q) * nv;
#pragma unroll
            for (int j = 0; j < kResMax; j++)
              if (j < nv) acc[j] -= ck[j] * v[q];
          }
      }
    }
#pragma unroll
    for (int j = 0; j < kResMax; j++)
      if (j < nv) t.p[j][i] = acc[j];
  }
}

__global__ void __launch_bounds__(kBT) k_axpy_s(double *__restrict__ y, const double *__restrict__ x, double a,
                                                size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) y[i] += a * x[i];
}

// dst = s * src, stored as TO
template <typename TO>
__global__ void __launch_bounds__(kBT) k_store(TO *__restrict__ dst, const double *__restrict__ src, double s,
                                               size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = (TO)(s * src[i]);
}

// v[i] += s * u_i with u_i uniform in [-1, 1) from a counter hash of
// (seed, i) (the splitmix64 finalizer), reproducible for any grid
__global__ void __launch_bounds__(kBT) k_add_noise(double *__restrict__ v, size_t n, unsigned long long seed,
                                                   double s) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    unsigned long long z = seed + 0x9e3779b97f4a7c15ull * (unsigned long long)(i + 1);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    v[i] += s * ((double)(z >> 11) * 0x1.0p-52 - 1.0);
  }
}

__global__ void __launch_bounds__(kBT) k_zero(double *__restrict__ p, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0.0;
}

// Cholesky S = L L^T of an n x n column-major S (ld); false when S is not
// numerically positive definite.
bool host_cholesky(int n, const double *S, int ld, std::vector<double> &L) {
  L.assign((size_t)n * n, 0.0);
  for (int j = 0; j < n; j++) {
    double s = S[j + (size_t)ld * j];
    for (int k = 0; k < j; k++) s -= L[j + (size_t)n * k] * L[j + (size_t)n * k];
    if (!(s > 0.0)) return false;
    const double l = std::sqrt(s);
    L[j + (size_t)n * j] = l;
    for (int i = j + 1; i < n; i++) {
      double t = S[i + (size_t)ld * j];
      for (int k = 0; k < j; k++) t -= L[i + (size_t)n * k] * L[j + (size_t)n * k];
      L[i + (size_t)n * j] = t / l;
    }
  }
  return true;
}

}  // namespace

// Generalized symmetric eigenproblem H y = theta S y of order n (column-
// major, ld), S positive definite: S = L L^T, A = L^-1 H L^-T, A = Q diag(E)
// Q^T (host_syev), Y = L^-T Q (S-orthonormal columns), ascending.
void host_sygv(int n, const double *H, const double *S, int ld, double *Y, double *E) {
  std::vector<double> L;
  if (!host_cholesky(n, S, ld, L)) {
    std::string diag;
    for (int i = 0; i < n && i < 24; i++) diag += " " + std::to_string(S[i + (size_t)ld * i]);
    throw Error(3, "block Davidson: basis overlap matrix (order " + std::to_string(n) +
                       ") is not positive definite; diagonal:" + diag);
  }
  // A = L^-1 H L^-T: first B = L^-1 H (forward substitution per column),
  // then A = L^-1 B^T (H symmetric, so (L^-1 H L^-T)^T = L^-1 (L^-1 H)^T)
  std::vector<double> B((size_t)n * n), A((size_t)n * n), Q((size_t)n * n);
  auto fwd = [&](const double *rhs, int rld, double *out) {  // out = L^-1 rhs (n x n)
    for (int c = 0; c < n; c++)
      for (int i = 0; i < n; i++) {
        double t = rhs[i + (size_t)rld * c];
        for (int k = 0; k < i; k++) t -= L[i + (size_t)n * k] * out[k + (size_t)n * c];
        out[i + (size_t)n * c] = t / L[i + (size_t)n * i];
      }
  };
  fwd(H, ld, B.data());
  std::vector<double> Bt((size_t)n * n);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) Bt[j + (size_t)n * i] = B[i + (size_t)n * j];
  fwd(Bt.data(), n, A.data());
  for (int i = 0; i < n; i++)  // symmetrize the rounding
    for (int j = 0; j < i; j++) {
      const double v = 0.5 * (A[i + (size_t)n * j] + A[j + (size_t)n * i]);
      A[i + (size_t)n * j] = A[j + (size_t)n * i] = v;
    }
  host_syev(n, A.data(), Q.data(), E);
  // Y = L^-T Q (back substitution per column)
  for (int c = 0; c < n; c++)
    for (int i = n - 1; i >= 0; i--) {
      double t = Q[i + (size_t)n * c];
      for (int k = i + 1; k < n; k++) t -= L[k + (size_t)n * i] * Y[k + (size_t)n * c];
      Y[i + (size_t)n * c] = t / L[i + (size_t)n * i];
    }
}

int davidson_block(Context &c, int nroots, int block, int max_it, double eps, double max_time, bool mixed,
                   const double *W0, double *Wout, double *Eout, int *n_products) {
  if (nroots < 1 || block < nroots || block > kResMax)
    throw Error(1, "davidson_block: need 1 <= nroots <= block <= 16");
  const size_t N = c.N;
  const dim3 g(kBG), bt(kBT);
  // spin-flip sectors (equal string sets): every column is even (X = X^T)
  // or odd (X = -X^T) and its product is the symmetric one of its sector;
  // H and S are block-diagonal by sector, so the Rayleigh-Ritz problem is
  // solved per sector and every Ritz vector belongs to one.  Mode 1: even
  // sector, 2: odd sector, 3: both in one basis: the lowest roots of H, and
  // the lowest pair of a sector that holds none of them is converged too
  // (and corrected every iteration), so that neither sector is left
  // unexplored.
  const int symmode = c.symmetric;
  if (symmode && !sigma_sym_available(c))
    throw Error(1, "spin-flip sectors need equal alpha and beta string sets and room for the dense blocks");
  const int nrep = symmode == 3 ? 1 : 0;  // at most one sector holds no root
  const int nslot = block + nrep;  // residual slots
  if (nslot > kResMax) throw Error(1, "davidson_block: with both sectors, block must be below 16");
  // ---- basis size and storage
  const int r64cap = mixed ? nroots + nrep : kNbMax;  // FP64 columns (all of them in FP64 mode)
  auto keep_of = [&](int m) { return std::min({kOutMax, m - nslot, std::max(nroots + nslot, m / 3)}); };
  const int m_min = nroots + nrep + 2 * nslot;
  auto need = [&](int m) {
    const int r64 = std::min(m, r64cap), nf = m - r64;
    return (double)N * (8.0 * (2.0 * r64 + nslot) + 4.0 * 2.0 * nf);
  };
  int m = c.blk_m_req > 0 ? c.blk_m_req : kNbMax;
  if (m < m_min || m > kNbMax)
    throw Error(1, "davidson_block: basis size must lie in [nroots + 2*block (+3 with both sectors), 64] (" +
                       std::to_string(m_min) + " here)");
  // buffers of an earlier solve are reused when they were sized for the
  // same request (mode, FP64 column count, block, requested basis size)
  const bool same_layout = c.blk_C64 && c.blk_mixed == (int)mixed && c.blk_r64cap == r64cap &&
                           c.blk_block >= nslot && c.blk_m_alloc_req == c.blk_m_req;
  if (!same_layout) {
    for (void *p : {(void *)c.blk_C64, (void *)c.blk_HC64, (void *)c.blk_C32, (void *)c.blk_HC32,
                    (void *)c.blk_R})
      if (p) dfree(c, p);
    c.blk_C64 = c.blk_HC64 = c.blk_R = nullptr;
    c.blk_C32 = c.blk_HC32 = nullptr;
    c.blk_m = 0;
    size_t free_b = 0, total_b = 0;
    SBDX_HIP(hipMemGetInfo(&free_b, &total_b));
    const double budget = 0.8 * (double)free_b;
    if (c.blk_m_req <= 0)
      while (need(m) > budget && m - nslot >= m_min) m -= nslot;
    if (need(m) > budget)
      throw Error(4, "davidson_block: a basis of " + std::to_string(m) + " vectors does not fit in device memory");
    const int r64 = std::min(m, r64cap), nf = m - r64;
    c.blk_C64 = (double *)dalloc(c, sizeof(double) * N * r64);
    c.blk_HC64 = (double *)dalloc(c, sizeof(double) * N * r64);
    if (nf > 0) {
      c.blk_C32 = (float *)dalloc(c, sizeof(float) * N * nf);
      c.blk_HC32 = (float *)dalloc(c, sizeof(float) * N * nf);
    }
    c.blk_R = (double *)dalloc(c, sizeof(double) * N * nslot);
    c.blk_m = m;
    c.blk_r64 = r64;
    c.blk_r64cap = r64cap;
    c.blk_mixed = mixed;
    c.blk_block = nslot;
    c.blk_m_alloc_req = c.blk_m_req;
  }
  m = c.blk_m;
  const int r64 = c.blk_r64;
  if (m < m_min) throw Error(1, "davidson_block: basis too small for this block");
  const int keep = keep_of(m);
  if (!c.blk_part) c.blk_part = (double *)dalloc(c, sizeof(double) * kBG * 2 * kNbMax * kResMax);
  if (!c.blk_small)
    c.blk_small = (double *)dalloc(c, sizeof(double) * (2 * kNbMax * kNbMax + kNbMax * kResMax + 4 * kNbMax));
  double *U = c.blk_small;                           // kNbMax x kNbMax, column-major: all Ritz pairs, ascending
  double *Ed = U + kNbMax * kNbMax;                  // their values
  double *Us = Ed + kNbMax;                          // selected pairs (residuals, restarts)
  double *Es = Us + kNbMax * kNbMax;                 // their values
  double *coef = Es + kNbMax;                        // CGS coefficients (d x nv)
  double *nrm = coef + kNbMax * kResMax;             // residual norms^2
  double *part = c.blk_part;
  c.blk_m_used = m;
  c.blk_keep_used = keep;
  std::vector<int> par(kNbMax, 0), rpar(kNbMax, 0);  // column parities, Ritz-pair parities
  struct SectorGuard {  // no graph replays here (columns alternate sectors); restore the caller's state
    Context &c;
    int graphs, spar;
    ~SectorGuard() {
      c.use_graphs = graphs;
      c.sym_par = spar;
    }
  } sguard{c, c.use_graphs, c.sym_par};
  c.use_graphs = 0;

  auto col_d = [&](double *b64, float *b32, int k) -> std::pair<double *, float *> {
    if (k < r64) return {b64 + (size_t)k * N, nullptr};
    return {nullptr, b32 + (size_t)(k - r64) * N};
  };
  auto cols = [&](bool prod, int k0, int k1) {  // basis (or product) columns [k0, k1)
    Cols L{};
    for (int k = k0; k < k1; k++) {
      auto p = col_d(prod ? c.blk_HC64 : c.blk_C64, prod ? c.blk_HC32 : c.blk_C32, k);
      if (p.first) L.d[L.nd++] = p.first;
      else L.f[L.nf++] = p.second;
    }
    return L;
  };
  auto outcols = [&](bool prod, int k0, int k1) {
    OutCols L{};
    for (int k = k0; k < k1; k++) {
      auto p = col_d(prod ? c.blk_HC64 : c.blk_C64, prod ? c.blk_HC32 : c.blk_C32, k);
      if (p.first) L.d[L.nd++] = p.first;
      else L.f[L.nf++] = p.second;
    }
    return L;
  };
  auto dcols = [&](const std::vector<double *> &v) {
    Cols L{};
    for (double *p : v) L.d[L.nd++] = p;
    return L;
  };
  auto t_start = std::chrono::steady_clock::now();
  auto elapsed = [&]() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  };
  // host copies of H and S (column-major, ld kNbMax)
  std::vector<double> H((size_t)kNbMax * kNbMax, 0.0), S((size_t)kNbMax * kNbMax, 0.0);
  std::vector<double> hU((size_t)kNbMax * kNbMax, 0.0), hE(kNbMax, 0.0), hR(kResMax, 0.0);
  std::vector<double> hUs((size_t)kNbMax * kNbMax, 0.0), hEs(kNbMax, 0.0);
  std::vector<double> gbuf(2 * kNbMax * kNbMax);
  // G = A^T B (na x nb, row-major a*nb + b) of two column lists, to dst
  // (device) and, when host is given, to the host
  auto gram = [&](const Cols &A, const Cols &B, double *dst, double *host) {
    const int na = A.nd + A.nf, nb = B.nd + B.nf, tot = na * nb;
    hipLaunchKernelGGL(k_bgram, g, bt, 0, c.stream, A, B, N, part);
    check_launch("k_bgram");
    hipLaunchKernelGGL(k_bsum, dim3((tot + 255) / 256), dim3(256), 0, c.stream, tot, kBG, part, dst);
    check_launch("k_bsum");
    if (host) {
      SBDX_HIP(hipMemcpyAsync(host, dst, sizeof(double) * tot, hipMemcpyDeviceToHost, c.stream));
      SBDX_HIP(hipStreamSynchronize(c.stream));
    }
  };
  // columns j in [d0, d1) of H and S: H[i][j] = HC_min(i,j) . C_max(i,j),
  // S[i][j] = C_i . C_j, mirrored
  auto columns = [&](int d0, int d1) {
    ProfScope ps(c, "blk_subspace");
    for (int j0 = d0; j0 < d1; j0 += kResMax) {
      const int j1 = std::min(d1, j0 + kResMax), nb = j1 - j0;
      // two grams, each over at most 64 x 16 tiles: HC[0,d1) x C[j0,j1), C[0,d1) x C[j0,j1)
      const Cols Bc = cols(false, j0, j1);
      gram(cols(true, 0, d1), Bc, coef, gbuf.data());
      for (int i = 0; i < d1; i++)
        for (int jj = 0; jj < nb; jj++) {
          const int j = j0 + jj;
          const double v = par[i] * par[j] < 0 ? 0.0 : gbuf[(size_t)i * nb + jj];  // HC_i . C_j
          if (i <= j) {
            H[i + (size_t)kNbMax * j] = v;
            H[j + (size_t)kNbMax * i] = v;
          }
        }
      gram(cols(false, 0, d1), Bc, coef, gbuf.data());
      for (int i = 0; i < d1; i++)
        for (int jj = 0; jj < nb; jj++) {
          const int j = j0 + jj;
          const double v = par[i] * par[j] < 0 ? 0.0 : gbuf[(size_t)i * nb + jj];
          if (i <= j) {
            S[i + (size_t)kNbMax * j] = v;
            S[j + (size_t)kNbMax * i] = v;
          }
        }
    }
  };
  // Rayleigh-Ritz on the d-column basis: U, E on the device and host
  auto rayleigh_ritz = [&](int d) {
    // per sector (one group when sectors are off): the group's columns,
    // its pairs, merged in ascending order (ties: even first)
    std::vector<std::pair<double, std::pair<int, std::vector<double>>>> pairs;
    for (int sgn : {1, -1, 0}) {
      std::vector<int> idx;
      for (int k = 0; k < d; k++)
        if (par[k] == sgn) idx.push_back(k);
      const int q = (int)idx.size();
      if (!q) continue;
      std::vector<double> Y((size_t)q * q), Hs((size_t)q * q), Ss((size_t)q * q), Eq(q);
      for (int b = 0; b < q; b++)
        for (int a = 0; a < q; a++) {
          Hs[a + (size_t)q * b] = H[idx[a] + (size_t)kNbMax * idx[b]];
          Ss[a + (size_t)q * b] = S[idx[a] + (size_t)kNbMax * idx[b]];
        }
      host_sygv(q, Hs.data(), Ss.data(), q, Y.data(), Eq.data());
      for (int b = 0; b < q; b++) {
        std::vector<double> u(d, 0.0);
        for (int a = 0; a < q; a++) u[idx[a]] = Y[a + (size_t)q * b];
        pairs.push_back({Eq[b], {sgn, std::move(u)}});
      }
    }
    std::stable_sort(pairs.begin(), pairs.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    std::fill(hU.begin(), hU.end(), 0.0);
    for (int j = 0; j < d; j++) {
      hE[j] = pairs[j].first;
      rpar[j] = pairs[j].second.first;
      for (int i = 0; i < d; i++) hU[i + (size_t)kNbMax * j] = pairs[j].second.second[i];
    }
    SBDX_HIP(hipMemcpyAsync(U, hU.data(), sizeof(double) * kNbMax * d, hipMemcpyHostToDevice, c.stream));
    SBDX_HIP(hipMemcpyAsync(Ed, hE.data(), sizeof(double) * d, hipMemcpyHostToDevice, c.stream));
    SBDX_HIP(hipStreamSynchronize(c.stream));  // hU and hE are rewritten by the next call
  };
  // Us, Es: the Ritz pairs L[0..) of the last Rayleigh-Ritz step, in order
  auto upload_sel = [&](const std::vector<int> &L, int d) {
    std::fill(hUs.begin(), hUs.end(), 0.0);
    for (size_t q = 0; q < L.size(); q++) {
      hEs[q] = hE[L[q]];
      for (int i = 0; i < d; i++) hUs[i + (size_t)kNbMax * q] = hU[i + (size_t)kNbMax * L[q]];
    }
    SBDX_HIP(hipMemcpyAsync(Us, hUs.data(), sizeof(double) * kNbMax * L.size(), hipMemcpyHostToDevice, c.stream));
    SBDX_HIP(hipMemcpyAsync(Es, hEs.data(), sizeof(double) * L.size(), hipMemcpyHostToDevice, c.stream));
    SBDX_HIP(hipStreamSynchronize(c.stream));
  };
  // the pairs that must converge: the nroots lowest, and with both sectors
  // the lowest pair of a sector that holds none of them
  auto required = [&](int d) {
    std::vector<int> req;
    for (int j = 0; j < std::min(nroots, d); j++) req.push_back(j);
    if (symmode == 3)
      for (int sgn : {1, -1}) {
        bool has = false;
        for (int j : req) has = has || rpar[j] == sgn;
        if (has) continue;
        for (int j = 0; j < d; j++)
          if (rpar[j] == sgn) {
            req.push_back(j);
            break;
          }
      }
    return req;
  };
  auto product = [&](int k0, int k1) {  // HC[k0,k1) from C[k0,k1)
    if (k1 <= k0) return;
    if (symmode) {  // one symmetric product per column, in its sector
      for (int k = k0; k < k1; k++) {
        c.sym_par = par[k];
        if (k < r64) {
          ProfScope ps(c, "blk_sigma_sym");
          sigma_sym(c, c.blk_C64 + (size_t)k * N, c.blk_HC64 + (size_t)k * N);
        } else {
          ProfScope ps(c, "blk_sigma_sym_f32");
          sigma_sym_f32_raw(c, c.blk_C32 + (size_t)(k - r64) * N, c.blk_HC32 + (size_t)(k - r64) * N);
        }
      }
      return;
    }
    const int a = std::min(k0, r64), b = std::min(k1, r64);
    if (b > a) {
      ProfScope ps(c, "blk_sigma");
      sigma_block(c, b - a, c.blk_C64 + (size_t)a * N, c.blk_HC64 + (size_t)a * N);
    }
    const int fa = std::max(k0, r64) - r64, fb = std::max(k1, r64) - r64;
    if (fb > fa) {
      // FP32 storage: in mixed mode these are correction columns, whose
      // products are FP32 by design; in FP64 mode there are none
      ProfScope ps(c, "blk_sigma_f32");
      sigma_block_f32_raw(c, fb - fa, c.blk_C32 + (size_t)fa * N, c.blk_HC32 + (size_t)fa * N);
    }
  };
  // store FP64 vector v (scaled by s) as basis column k
  auto store = [&](int k, const double *v, double s) {
    auto p = col_d(c.blk_C64, c.blk_C32, k);
    if (p.first)
      hipLaunchKernelGGL(k_store<double>, g, bt, 0, c.stream, p.first, v, s, N);
    else
      hipLaunchKernelGGL(k_store<float>, g, bt, 0, c.stream, p.second, v, s, N);
    check_launch("k_store");
  };
  // modified Gram-Schmidt of the FP64 vectors vs among themselves; a
  // vector whose squared norm falls to 1e-20 of ref[j] or to zero is
  // dropped; returns the kept ones with their norms (not yet scaled)
  auto mgs = [&](const std::vector<double *> &vs, const std::vector<double> &ref, std::vector<double *> &kept,
                 std::vector<double> &knorm, std::vector<int> &kidx) {
    kept.clear();
    knorm.clear();
    kidx.clear();
    for (size_t j = 0; j < vs.size(); j++) {
      double *v = vs[j];
      for (size_t k = 0; k < kept.size(); k++) {
        const double o = device_dot(c, kept[k], v, N) / (knorm[k] * knorm[k]);
        hipLaunchKernelGGL(k_axpy_s, g, bt, 0, c.stream, v, kept[k], -o, N);
        check_launch("k_axpy_s");
      }
      const double n2 = device_dot(c, v, v, N);
      if (!(n2 > 1e-20 * ref[j]) || !(n2 > 0.0)) continue;
      kept.push_back(v);
      knorm.push_back(std::sqrt(n2));
      kidx.push_back((int)j);
    }
  };

  // ---- coarse start: the lowest roots of the problem on the first
  // coarse_m strings of each set, found by this solver (nested levels, at
  // most kCoarseIt iterations), lifted back as the start vectors (zero
  // elsewhere; the basis sign depends only on the two strings).  The
  // sub-problem starts from the given vectors restricted to it when each
  // keeps at least 1e-3 of its squared norm there, else from unit vectors on
  // its lowest diagonal elements; when it cannot start, the given vectors
  // are used as they are.
  const double *start = W0;
  std::unique_ptr<TmpBuf> lifted;
  c.coarse_products = 0;
  c.coarse_energy = 0.0;
  c.coarse_seconds = 0.0;
  const size_t mc = c.coarse_m;
  if (mc > 0 && mc < c.na && mc < c.nb) {
    ProfScope ps(c, "coarse_start");
    const auto tc = std::chrono::steady_clock::now();
    Context *sub = make_sub_context(c, mc);
    struct Owner {
      Context *p;
      ~Owner() { destroy_context(p); }
    } owner{sub};
    sub->precision = c.precision;
    if (symmode && sigma_sym_available(*sub)) sub->symmetric = symmode;
    sub->coarse_m = mc / 4 >= 256 ? mc / 4 : 0;
    sub->coarse_eps = c.coarse_eps;
    const size_t Ns = mc * mc;
    TmpBuf sw0(sizeof(double) * Ns * nroots), sw1(sizeof(double) * Ns * nroots);
    double *s0 = (double *)sw0.p, *s1 = (double *)sw1.p;
    SBDX_HIP(hipStreamSynchronize(c.stream));
    for (int j = 0; j < nroots; j++)
      SBDX_HIP(hipMemcpy2DAsync(s0 + j * Ns, sizeof(double) * mc, W0 + j * N, sizeof(double) * c.nb,
                                sizeof(double) * mc, mc, hipMemcpyDeviceToDevice, sub->stream));
    SBDX_HIP(hipStreamSynchronize(sub->stream));
    bool restricted_ok = true;
    for (int j = 0; j < nroots && restricted_ok; j++) {
      const double full = device_dot(c, W0 + j * N, W0 + j * N, N);
      const double pt = device_dot(*sub, s0 + j * Ns, s0 + j * Ns, Ns);
      restricted_ok = pt > 1e-3 * full && pt > 0.0;
    }
    if (!restricted_ok) {
      std::vector<double> hh(Ns);
      SBDX_HIP(hipMemcpyAsync(hh.data(), sub->hii, sizeof(double) * Ns, hipMemcpyDeviceToHost, sub->stream));
      SBDX_HIP(hipStreamSynchronize(sub->stream));
      std::vector<size_t> idx(Ns);
      for (size_t i = 0; i < Ns; i++) idx[i] = i;
      const size_t kk = std::min<size_t>(Ns, (size_t)nroots);
      std::partial_sort(idx.begin(), idx.begin() + kk, idx.end(),
                        [&](size_t a, size_t b) { return hh[a] < hh[b] || (hh[a] == hh[b] && a < b); });
      std::vector<double> h0(Ns * nroots, 0.0);
      for (size_t k = 0; k < kk; k++) h0[k * Ns + idx[k]] = 1.0;
      SBDX_HIP(hipMemcpyAsync(s0, h0.data(), sizeof(double) * Ns * nroots, hipMemcpyHostToDevice, sub->stream));
      SBDX_HIP(hipStreamSynchronize(sub->stream));
    }
    std::vector<double> es(nroots);
    int np = 0;
    bool ok = true;
    try {
      // partly converged Ritz vectors already make far better starts than
      // unit vectors, and excited roots of the sub-problem can take long
      (void)davidson_block(*sub, nroots, block, std::min(max_it, kCoarseIt), c.coarse_eps, max_time, mixed, s0,
                           s1, es.data(), &np);
    } catch (const Error &) {
      ok = false;
    }
    SBDX_HIP(hipStreamSynchronize(sub->stream));
    if (ok) {
      lifted.reset(new TmpBuf(sizeof(double) * N * nroots));
      double *l = (double *)lifted->p;
      hipLaunchKernelGGL(k_zero, g, bt, 0, c.stream, l, N * nroots);
      check_launch("k_zero");
      for (int j = 0; j < nroots; j++)
        SBDX_HIP(hipMemcpy2DAsync(l + j * N, sizeof(double) * c.nb, s1 + j * Ns, sizeof(double) * mc,
                                  sizeof(double) * mc, mc, hipMemcpyDeviceToDevice, c.stream));
      SBDX_HIP(hipStreamSynchronize(c.stream));
      start = l;
      c.coarse_products = np;
      c.coarse_energy = es[0];
    }
    c.coarse_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - tc).count();
    if (c.verbose) {
      fprintf(stderr, "[sbdx] block coarse start on %zu x %zu strings: %s, E0 = %.12f, %d products, %.3f s\n", mc,
              mc, ok ? "used" : "skipped", c.coarse_energy, c.coarse_products, c.coarse_seconds);
      fflush(stderr);
    }
  }

  // ---- initial block: the start vectors, then unit vectors on the lowest
  // diagonal elements, each with a seeded random component carrying ~1% of
  // its squared norm, orthonormalized in the residual slots.  A start block
  // confined to one symmetry sector of H (spatial irreducible
  // representation, spin) never leaves it in exact arithmetic; the random
  // component gives the basis a share of every sector, so the solver finds
  // the lowest roots whatever symmetry they have.
  // With spin-flip sectors every start column is projected on one sector
  // (mode 3: the even and the odd part of each start vector, then extra
  // vectors alternately even and odd) and all of them get the random
  // component.  A projection that keeps less than 1e-6 of the start
  // vector's squared norm (the odd part of a closed-shell start) is
  // replaced by the unit pair on the lowest diagonal element of its sector
  // (odd: the lowest off-diagonal hii[a,b], as X[a,b] = -X[b,a] = 1).
  std::vector<double *> Rs(nslot);
  for (int j = 0; j < nslot; j++) Rs[j] = c.blk_R + (size_t)j * N;
  std::vector<int> src(block), spar(block, 0);  // source (start j, or -1-k: k-th extra), sector
  {
    int nextra = 0;
    for (int t = 0; t < block; t++) {
      if (symmode == 3) {
        spar[t] = t % 2 == 0 ? 1 : -1;
        src[t] = t < 2 * nroots ? t / 2 : -1 - nextra++;
      } else {
        spar[t] = symmode == 1 ? 1 : symmode == 2 ? -1 : 0;
        src[t] = t < nroots ? t : -1 - nextra++;
      }
    }
    // the block + 1 lowest diagonal elements, over all determinants and
    // over those off the string-exchange diagonal (a != b), by one host pass
    std::vector<size_t> low_all, low_off;
    if (nextra > 0 || symmode) {
      std::vector<double> hh(N);
      SBDX_HIP(hipMemcpyAsync(hh.data(), c.hii, sizeof(double) * N, hipMemcpyDeviceToHost, c.stream));
      SBDX_HIP(hipStreamSynchronize(c.stream));
      const size_t kk = (size_t)block + 1, nbs = c.nb;
      auto less = [&](size_t a, size_t b) { return hh[a] < hh[b] || (hh[a] == hh[b] && a < b); };
      auto push = [&](std::vector<size_t> &v, size_t i) {
        if (v.size() == kk && !less(i, v.back())) return;
        v.insert(std::upper_bound(v.begin(), v.end(), i, less), i);
        if (v.size() > kk) v.pop_back();
      };
      for (size_t i = 0; i < N; i++) {
        push(low_all, i);
        if (i / nbs != i % nbs) push(low_off, i);
      }
    }
    // |u|^2 ~ N/3 for uniform u in [-1, 1): scale to ~0.01 squared norm
    const double sn = 0.1 / std::sqrt((double)N / 3.0);
    const size_t nbs = c.nb;
    auto unit = [&](double *v, int p, size_t k) {  // unit vector (pair) of sector p on the k-th lowest element
      hipLaunchKernelGGL(k_zero, g, bt, 0, c.stream, v, N);
      check_launch("k_zero");
      const std::vector<size_t> &L = p < 0 ? low_off : low_all;
      if (k >= L.size()) return;
      const size_t i = L[k], a = i / nbs, b = i % nbs;
      const double one = 1.0, sgn = p < 0 ? -1.0 : 1.0;
      SBDX_HIP(hipMemcpyAsync(v + i, &one, sizeof(double), hipMemcpyHostToDevice, c.stream));
      if (p && a != b) SBDX_HIP(hipMemcpyAsync(v + b * nbs + a, &sgn, sizeof(double), hipMemcpyHostToDevice, c.stream));
      SBDX_HIP(hipStreamSynchronize(c.stream));
    };
    int nrepl = 0;
    for (int t = 0; t < block; t++) {
      if (src[t] >= 0) {
        SBDX_HIP(hipMemcpyAsync(Rs[t], start + (size_t)src[t] * N, sizeof(double) * N, hipMemcpyDeviceToDevice,
                                c.stream));
        if (spar[t]) {
          const double n0 = device_dot(c, Rs[t], Rs[t], N);
          symmetrize(c, Rs[t], spar[t]);
          if (!(device_dot(c, Rs[t], Rs[t], N) > 1e-6 * n0)) unit(Rs[t], spar[t], (size_t)nrepl++);
        }
      } else {
        unit(Rs[t], spar[t], (size_t)(-1 - src[t]) + (symmode ? (size_t)nrepl : 0));
      }
      if (src[t] < 0 || symmode) {
        hipLaunchKernelGGL(k_add_noise, g, bt, 0, c.stream, Rs[t], N, 0x5bd1e995ull * (unsigned long long)(t + 1),
                           sn);
        check_launch("k_add_noise");
      }
      if (spar[t]) symmetrize(c, Rs[t], spar[t]);
    }
    SBDX_HIP(hipStreamSynchronize(c.stream));
  }
  int d = 0;
  {
    std::vector<double> ref(block);
    for (int j = 0; j < block; j++) ref[j] = device_dot(c, Rs[j], Rs[j], N);
    std::vector<double *> kept, first(Rs.begin(), Rs.begin() + block);
    std::vector<double> kn;
    std::vector<int> ki;
    mgs(first, ref, kept, kn, ki);
    if ((int)kept.size() < nroots) throw Error(1, "davidson_block: start vectors are linearly dependent");
    for (size_t k = 0; k < kept.size(); k++) {
      par[k] = spar[ki[k]];
      store((int)k, kept[k], 1.0 / kn[k]);
    }
    d = (int)kept.size();
  }
  product(0, d);
  int products = d;
  columns(0, d);
  bool done = false, exact_now = !mixed || d <= r64;
  c.trace.clear();
  c.blk_E.assign(nroots, 0.0);
  c.blk_resid.assign(nroots, 0.0);
  c.blk_restarts = 0;
  std::vector<double> rhist;
  int diag_left = 0;
  c.mlp.used_mlp = c.mlp.used_diag = 0;
  std::unique_ptr<TmpBuf> mlp_tmp;
  const double eps_reg = 1.0e-12;
  for (int it = 0;; it++) {
    // the learned preconditioner unless the residual stalled (below)
    const bool use_mlp_now = c.precond == 1 && c.mlp.valid && diag_left == 0;
    std::vector<int> req, P;  // required pairs; pairs whose residuals are formed (ascending)
    int nres = 0;
    {
      ProfScope ps(c, "blk_ritz");
      rayleigh_ritz(d);
      req = required(d);
      P = req;
      const int cap = std::max((int)req.size(), std::min(d, block));
      for (int j = 0; j < d && (int)P.size() < cap; j++)
        if (std::find(P.begin(), P.end(), j) == P.end()) P.push_back(j);
      std::sort(P.begin(), P.end());
      nres = (int)P.size();
      upload_sel(P, d);
      OutD o{};
      for (int q = 0; q < nres; q++) o.p[q] = Rs[q];
      hipLaunchKernelGGL(k_bresid, g, bt, 0, c.stream, cols(false, 0, d), cols(true, 0, d), (const double *)Us,
                         kNbMax, (const double *)Es, nres, (const double *)c.hii, eps_reg, use_mlp_now ? 0 : 1, o,
                         N, part);
      check_launch("k_bresid");
      hipLaunchKernelGGL(k_bsum, dim3(1), dim3(256), 0, c.stream, nres, kBG, part, nrm);
      check_launch("k_bsum");
      SBDX_HIP(hipMemcpyAsync(hR.data(), nrm, sizeof(double) * nres, hipMemcpyDeviceToHost, c.stream));
      SBDX_HIP(hipStreamSynchronize(c.stream));
      for (int j = 0; j < nres; j++) hR[j] = std::sqrt(hR[j]);
    }
    // P is ascending and holds 0..nroots-1 first, so hR[j] is root j's
    double rmax = 0.0;
    auto resid_of = [&](int j) {  // +inf for a pair whose residual was not formed
      const auto it = std::find(P.begin(), P.end(), j);
      return it == P.end() ? HUGE_VAL : hR[it - P.begin()];
    };
    for (int j : req) rmax = std::max(rmax, resid_of(j));
    for (int j = 0; j < nroots; j++) {
      c.blk_E[j] = hE[j];
      c.blk_resid[j] = hR[j];
    }
    c.trace.push_back(hE[0]);
    c.trace.push_back(rmax);
    if (c.verbose) {
      fprintf(stderr, "[sbdx] block it %d dim %d products %d%s E0 = %.15f max|r| = %.3e |r_j|", it, d, products,
              exact_now ? " (exact)" : "", hE[0], rmax);
      for (int j : req) fprintf(stderr, " %.2e%s", resid_of(j), symmode == 3 ? (rpar[j] > 0 ? "e" : "o") : "");
      fprintf(stderr, "\n");
      fflush(stderr);
    }
    if (rmax < eps && exact_now) {
      done = true;
      break;
    }
    const bool out_of_time = elapsed() > max_time || it >= max_it;
    if (d + nslot > m || (rmax < eps && !exact_now) || out_of_time) {
      // ---- thick restart on the lowest Ritz vectors (and the required
      // ones), in place; to confirm convergence (or at the end) the
      // required pairs alone, so that every column and product is FP64 in
      // mixed mode
      ProfScope ps(c, "blk_restart");
      std::vector<int> K;
      if (out_of_time || rmax < eps) {
        K = req;
      } else {
        for (int j = 0; j < std::min(keep, d); j++) K.push_back(j);
        for (int j : req)
          if (std::find(K.begin(), K.end(), j) == K.end()) K.back() = j;  // past the roots, so never one of them
        std::sort(K.begin(), K.end());
      }
      const int nk = (int)K.size();
      upload_sel(K, d);
      {
        OutCols oc = outcols(false, 0, nk), oh = outcols(true, 0, nk);
        hipLaunchKernelGGL(k_bcomb, g, bt, 0, c.stream, cols(false, 0, d), (const double *)Us, kNbMax, oc, N);
        check_launch("k_bcomb");
        hipLaunchKernelGGL(k_bcomb, g, bt, 0, c.stream, cols(true, 0, d), (const double *)Us, kNbMax, oh, N);
        check_launch("k_bcomb");
      }
      // the restarted columns' sectors, before their exact products
      for (int j = 0; j < nk; j++) par[j] = rpar[K[j]];
      // exact FP64 products of the roots: on the confirming restart, and on
      // every restart once max|r| falls below 100 eps or below the level
      // where FP32 products' errors (1000 x 2^-24 |E|) can hold the
      // residuals up; before that the roots' products stay combinations of
      // the basis products
      double emax = 0.0;
      for (int j = 0; j < nroots; j++) emax = std::max(emax, std::fabs(hE[j]));
      // (every root, converged or not: a locked root's combined product
      // would otherwise collect the FP32 columns' errors restart after
      // restart)
      if (mixed && (nk <= r64 || rmax < std::max(100.0 * eps, kFp32Level * emax))) {
        const int ne = std::min(nk, r64);
        product(0, ne);
        products += ne;
      }
      d = nk;
      c.blk_restarts++;
      columns(0, d);
      exact_now = !mixed || d <= r64;
      if (out_of_time) {
        rayleigh_ritz(d);
        SBDX_HIP(hipStreamSynchronize(c.stream));
        for (int j = 0; j < nroots; j++) c.blk_E[j] = hE[j];
        break;
      }
      continue;  // Rayleigh-Ritz on the restarted basis gives the residuals
    }
    // ---- corrections for the unlocked pairs among the lowest nres; the
    // next 3 iterations use the diagonal step when max|r| has not dropped
    // 10% over the last 3 expansions
    rhist.push_back(rmax);
    const size_t nh = rhist.size();
    if (diag_left > 0) diag_left--;
    else if (nh >= 4 && rhist[nh - 1] > 0.9 * rhist[nh - 4]) diag_left = 3;
    std::vector<double *> sel;
    std::vector<int> selj;  // Ritz index of each correction
    std::vector<int> selq;  // its slot in P
    for (int q = 0; q < nres; q++) {
      const bool is_req = std::find(req.begin(), req.end(), P[q]) != req.end();
      if (!(is_req && hR[q] < eps)) {
        sel.push_back(Rs[q]);
        selj.push_back(P[q]);
        selq.push_back(q);
      }
    }
    if (use_mlp_now) {
      ProfScope ps(c, "precond_mlp");
      if (!mlp_tmp) mlp_tmp.reset(new TmpBuf(sizeof(double) * N));
      for (size_t q = 0; q < sel.size(); q++) {
        mlp_precond(c, sel[q], Es + selq[q], (double *)mlp_tmp->p);
        SBDX_HIP(hipMemcpyAsync(sel[q], mlp_tmp->p, sizeof(double) * N, hipMemcpyDeviceToDevice, c.stream));
        c.mlp.used_mlp++;
      }
    } else {
      c.mlp.used_diag += (int)sel.size();
    }
    int nv = 0;
    {
      ProfScope ps(c, "blk_orth");
      const int ns = (int)sel.size();
      std::vector<double> ref(ns);
      for (int q = 0; q < ns; q++) ref[q] = device_dot(c, sel[q], sel[q], N);
      OutD t{};
      for (int q = 0; q < ns; q++) t.p[q] = sel[q];
      for (int pass = 0; pass < 2; pass++) {  // classical Gram-Schmidt, twice
        const Cols X = cols(false, 0, d);
        gram(X, dcols(sel), coef, nullptr);
        hipLaunchKernelGGL(k_bsub, g, bt, 0, c.stream, X, (const double *)coef, ns, t, N);
        check_launch("k_bsub");
      }
      std::vector<double *> kept;
      std::vector<double> kn;
      std::vector<int> ki;
      mgs(sel, ref, kept, kn, ki);
      nv = (int)kept.size();
      for (int q = 0; q < nv; q++) {
        // each correction stays in its Ritz pair's sector; projecting it
        // there removes the rounding-level share of the other sector that
        // the Gram-Schmidt coefficients leave
        par[d + q] = rpar[selj[ki[q]]];
        if (par[d + q]) symmetrize(c, kept[q], par[d + q]);
        store(d + q, kept[q], 1.0 / kn[q]);
      }
    }
    if (nv == 0) throw Error(3, "davidson_block: no new direction (basis exhausted)");
    product(d, d + nv);
    products += nv;
    columns(d, d + nv);
    d += nv;
    exact_now = !mixed;
  }
  // ---- outputs: the root Ritz vectors of the final basis
  {
    OutCols o{};
    for (int j = 0; j < nroots; j++) o.d[o.nd++] = Wout + (size_t)j * N;
    hipLaunchKernelGGL(k_bcomb, g, bt, 0, c.stream, cols(false, 0, d), (const double *)U, kNbMax, o, N);
    check_launch("k_bcomb");
  }
  SBDX_HIP(hipStreamSynchronize(c.stream));
  for (int j = 0; j < nroots; j++) Eout[j] = c.blk_E[j];
  c.blk_par.assign(rpar.begin(), rpar.begin() + nroots);
  if (n_products) *n_products = products;
  return done ? 0 : 5;
}

}  // namespace sbdx
