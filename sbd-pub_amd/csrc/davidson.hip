// Single-root Davidson on the device, following chemistry/tpb/davidson.h:667-930
// (and chemistry/basic/davidson_thrust.h:85-373): per step one sigma, the new
// column of the subspace matrix, a Rayleigh-Ritz eigensolve, W and the
// residual R = E0*W - H*W, both normalized; stop when |R| < eps; otherwise
// the diagonal preconditioner R/(E0-d) (1e-12 guard), modified Gram-Schmidt
// and normalization give the next basis vector; restart from W every nb.
//
// Every reduction is a fixed two-level tree (per-block partial sums in a
// fixed grid, then a fixed-order sum), so results are reproducible run to
// run.  The subspace eigenproblem (at most 64 x 64, LAPACK dsyev in the
// reference, framework/hp_numeric.h:503-522) is solved on the host by
// Householder tridiagonalization and implicit QR, through pinned buffers.
#include <rocblas/rocblas.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <string>

#include "common.h"

namespace sbdx {

constexpr int kRedBlocks = 1024;
// FP32 noise level of a product relative to |E0| (1000 x 2^-24): below
// kFp32Floor |E0| a stalled mixed-precision solve switches to FP64 products
constexpr double kFp32Floor = 1000.0 * 5.9604644775390625e-8;
// distance above another sector's result, relative to |E0|, beyond which an
// abandonable stage keeps its anchor products in FP32
constexpr double kAnchorMargin = 1e-3;
constexpr int kRedThreads = 256;

struct OutList {
  double *p[8];
};

// Basis columns for the streaming kernels: column k is p[k] (FP64) or, when
// bit k of f32 is set, pf[k] (FP32 storage of an FP32 product, widened on
// load: the values are those of the FP64 copy it replaces).
struct VecList {
  const double *p[kNbMax];
  const float *pf[kNbMax];
  uint64_t f32;
};
__device__ __forceinline__ double vget(const VecList &X, int k, size_t i) {
  return ((X.f32 >> k) & 1ull) ? (double)X.pf[k][i] : X.p[k][i];
}

__device__ __forceinline__ double block_sum(double v, double *sh) {
  // deterministic LDS tree over 256 threads
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = kRedThreads / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// part[b*nv + k] = sum over this block's slice of X_k . y.  The vectors go
// in groups of 8 sharing one read of y (y is read ceil(nv/8) times, each X_k
// once); every thread sums its elements in index order, then a fixed tree.
// gate (may be null): the kernel does nothing when *gate == 0 (the
// conditional second Gram-Schmidt pass, decided on the device).
__global__ void __launch_bounds__(kRedThreads) k_dots(VecList X, int nv, const double *__restrict__ y,
                                                      size_t n, double *__restrict__ part,
                                                      const int *__restrict__ gate) {
  if (gate && *gate == 0) return;  // uniform per block, before any barrier
  __shared__ double sh[kRedThreads];
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int k0 = 0; k0 < nv; k0 += 8) {
    const int m = nv - k0 < 8 ? nv - k0 : 8;
    double s[8];
#pragma unroll
    for (int j = 0; j < 8; j++) s[j] = 0.0;
    for (size_t i = i0; i < n; i += stride) {
      const double yi = y[i];
#pragma unroll
      for (int j = 0; j < 8; j++)
        if (j < m) s[j] += vget(X, k0 + j, i) * yi;
    }
    for (int j = 0; j < m; j++) {
      const double r = block_sum(s[j], sh);
      if (threadIdx.x == 0) part[(size_t)blockIdx.x * nv + k0 + j] = r;
    }
  }
}

// Fixed-order sum of part[b*stride + off] over b < nblk by one whole
// wavefront: lane l adds b = l, l + 64, ... in order, then a butterfly over
// the 64 lanes (each step adds a pair of lanes' values in both lanes, so
// every lane ends with the same total).  A single thread summing the 1024
// partials one load at a time took ~0.1 ms.
__device__ __forceinline__ double wave_sum(const double *part, int nblk, int stride, int off) {
  const int lane = threadIdx.x & 63;
  double s = 0.0;
  for (int b = lane; b < nblk; b += 64) s += part[(size_t)b * stride + off];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  return s;
}
// threads of the one-block finishing kernels (4 wavefronts)
constexpr int kFinishThreads = 256;

// Fixed-order sum of the partials of column k; written to H[k + ld*col] and
// H[col + ld*k] (the symmetric subspace matrix, column-major).  Wavefront w
// sums columns w, w + 4, ...
__global__ void __launch_bounds__(kFinishThreads) k_finish_hcol(int nv, int nblk, const double *__restrict__ part,
                                                                double *H, int ld, int col) {
  for (int k = threadIdx.x >> 6; k < nv; k += kFinishThreads / 64) {
    const double s = wave_sum(part, nblk, nv, k);
    if ((threadIdx.x & 63) == 0) {
      H[k + (size_t)ld * col] = s;
      H[col + (size_t)ld * k] = s;
    }
  }
}

// ---- subspace eigensolver (host): eigenpairs of the leading n x n block of
// the symmetric subspace matrix, ascending.  Householder reduction to
// tridiagonal form with the reflectors accumulated into Q (Golub & Van Loan,
// Matrix Computations, Alg. 8.3.1), then implicit symmetric QR steps with a
// Wilkinson shift on the unreduced trailing block (Alg. 8.3.2-8.3.3), Givens
// rotations applied to T and Q, deflating off-diagonals below
// 1e-16 (|t_ii| + |t_i+1,i+1|).  O(n^3) for n <= 64: well under a
// millisecond, where one thread on the device took ~100 ms at n = 40.
// A, Q: n x n column-major (leading dimension n); E: n values.
void host_syev(int n, double *A, double *Q, double *E) {
  auto a = [&](int i, int j) -> double & { return A[i + (size_t)n * j]; };
  auto q = [&](int i, int j) -> double & { return Q[i + (size_t)n * j]; };
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) q(i, j) = i == j ? 1.0 : 0.0;
  std::vector<double> v(n), p(n), w(n);
  for (int k = 0; k + 2 < n; k++) {
    // reflector P = I - beta v v^T with P x = alpha e1, x = A[k+1:, k]
    double sig = 0.0;
    for (int i = k + 2; i < n; i++) sig += a(i, k) * a(i, k);
    const double x0 = a(k + 1, k);
    if (sig == 0.0) continue;
    const double mu = std::sqrt(x0 * x0 + sig);
    const double v0 = x0 <= 0.0 ? x0 - mu : -sig / (x0 + mu);
    const double beta = 2.0 * v0 * v0 / (sig + v0 * v0);
    for (int i = 0; i < n; i++) v[i] = 0.0;
    v[k + 1] = 1.0;
    for (int i = k + 2; i < n; i++) v[i] = a(i, k) / v0;
    // A <- P A P on rows/columns k+1.. (and column/row k)
    for (int i = k; i < n; i++) {
      double t = 0.0;
      for (int j = k + 1; j < n; j++) t += a(i, j) * v[j];
      p[i] = beta * t;
    }
    double pv = 0.0;
    for (int i = k + 1; i < n; i++) pv += p[i] * v[i];
    for (int i = k + 1; i < n; i++) w[i] = p[i] - 0.5 * beta * pv * v[i];
    for (int j = k + 1; j < n; j++)
      for (int i = k + 1; i < n; i++) a(i, j) -= v[i] * w[j] + w[i] * v[j];
    // with this choice of v0 the reflector maps x to |x| e1 (Alg. 5.1.1)
    a(k + 1, k) = a(k, k + 1) = mu;
    for (int i = k + 2; i < n; i++) a(i, k) = a(k, i) = 0.0;
    // Q <- Q P
    for (int i = 0; i < n; i++) {
      double t = 0.0;
      for (int j = k + 1; j < n; j++) t += q(i, j) * v[j];
      t *= beta;
      for (int j = k + 1; j < n; j++) q(i, j) -= t * v[j];
    }
  }
  // implicit symmetric QR on the tridiagonal T (held in A)
  std::vector<double> d(n), e(n > 1 ? n - 1 : 1);
  for (int i = 0; i < n; i++) d[i] = a(i, i);
  for (int i = 0; i + 1 < n; i++) e[i] = a(i + 1, i);
  const double tiny = 1e-16;
  for (int iter = 0; iter < 60 * n; iter++) {
    for (int i = 0; i + 1 < n; i++)
      if (std::fabs(e[i]) <= tiny * (std::fabs(d[i]) + std::fabs(d[i + 1]))) e[i] = 0.0;
    int hi = n - 1;
    while (hi > 0 && e[hi - 1] == 0.0) hi--;
    if (hi == 0) break;
    int lo = hi - 1;
    while (lo > 0 && e[lo - 1] != 0.0) lo--;
    // Wilkinson shift from the trailing 2x2 of the block [lo, hi]
    const double dd = 0.5 * (d[hi - 1] - d[hi]);
    const double b2 = e[hi - 1] * e[hi - 1];
    const double den = dd + (dd >= 0.0 ? 1.0 : -1.0) * std::sqrt(dd * dd + b2);
    const double mu = den != 0.0 ? d[hi] - b2 / den : d[hi];
    double x = d[lo] - mu, z = e[lo];
    for (int k = lo; k < hi; k++) {
      // Givens G = [c s; -s c] with G^T [x; z] = [r; 0]
      const double r = std::hypot(x, z);
      const double cs = r != 0.0 ? x / r : 1.0, sn = r != 0.0 ? -z / r : 0.0;
      if (k > lo) e[k - 1] = r;
      // T <- G^T T G on rows/columns k, k+1
      const double dk = d[k], dk1 = d[k + 1], ek = e[k];
      d[k] = cs * cs * dk - 2.0 * cs * sn * ek + sn * sn * dk1;
      d[k + 1] = sn * sn * dk + 2.0 * cs * sn * ek + cs * cs * dk1;
      e[k] = cs * sn * (dk - dk1) + (cs * cs - sn * sn) * ek;
      if (k + 1 < hi) {
        z = -sn * e[k + 1];
        e[k + 1] *= cs;
      }
      x = e[k];
      for (int i = 0; i < n; i++) {
        const double qk = q(i, k), qk1 = q(i, k + 1);
        q(i, k) = cs * qk - sn * qk1;
        q(i, k + 1) = sn * qk + cs * qk1;
      }
    }
  }
  // ascending order
  std::vector<int> idx(n);
  for (int i = 0; i < n; i++) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int x1, int x2) { return d[x1] < d[x2]; });
  std::vector<double> Qs((size_t)n * n);
  for (int j = 0; j < n; j++) {
    E[j] = d[idx[j]];
    for (int i = 0; i < n; i++) Qs[i + (size_t)n * j] = q(i, idx[j]);
  }
  std::copy(Qs.begin(), Qs.end(), Q);
}

// Eigenpairs of the leading n x n block of the device subspace matrix H
// (column-major, ld): to the host (pinned staging), host_syev, U (ld) and E
// back to the device.  Synchronous.
static void subspace_eig(Context &c, int n, int ld, const double *H, double *U, double *E) {
  if (!c.hpin) SBDX_HIP(hipHostMalloc((void **)&c.hpin, sizeof(double) * (3 * kNbMax * kNbMax + kNbMax)));
  double *hA = c.hpin, *hQ = c.hpin + kNbMax * kNbMax, *hU = c.hpin + 2 * kNbMax * kNbMax;
  double *hE = c.hpin + 3 * kNbMax * kNbMax;
  SBDX_HIP(hipMemcpy2DAsync(hA, sizeof(double) * n, H, sizeof(double) * ld, sizeof(double) * n, n,
                            hipMemcpyDeviceToHost, c.stream));
  SBDX_HIP(hipStreamSynchronize(c.stream));
  host_syev(n, hA, hQ, hE);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < ld; i++) hU[i + (size_t)ld * j] = i < n ? hQ[i + (size_t)n * j] : 0.0;
  SBDX_HIP(hipMemcpyAsync(U, hU, sizeof(double) * ld * n, hipMemcpyHostToDevice, c.stream));
  SBDX_HIP(hipMemcpyAsync(E, hE, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
  SBDX_HIP(hipStreamSynchronize(c.stream));
}

// W = sum_k U[k] C_k ; R = -sum_k U[k] HC_k + E0*W ; partial |W|^2, |R|^2.
__global__ void __launch_bounds__(kRedThreads) k_ritz(VecList Cv, VecList HCv, int nv,
                                                      const double *__restrict__ U,
                                                      const double *__restrict__ E, size_t n,
                                                      double *__restrict__ W, double *__restrict__ R,
                                                      double *__restrict__ part) {
  __shared__ double sh[kRedThreads];
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const double E0 = E[0];
  double sw = 0.0, sr = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    // vectors in groups of 8 whose loads issue together (index order of
    // the sums kept); lanes j >= m reload vector k0 and are not summed
    double w = 0.0, r = 0.0;
    for (int k0 = 0; k0 < nv; k0 += 8) {
      const int m = nv - k0 < 8 ? nv - k0 : 8;
      double cv[8], hv[8];
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int k = j < m ? k0 + j : k0;
        cv[j] = Cv.p[k][i];
        hv[j] = vget(HCv, k, i);
      }
#pragma unroll
      for (int j = 0; j < 8; j++)
        if (j < m) {
          w += cv[j] * U[k0 + j];
          r += hv[j] * (-U[k0 + j]);
        }
    }
    r += E0 * w;
    W[i] = w;
    R[i] = r;
    sw += w * w;
    sr += r * r;
  }
  const double a = block_sum(sw, sh);
  const double b = block_sum(sr, sh);
  if (threadIdx.x == 0) {
    part[2 * (size_t)blockIdx.x] = a;
    part[2 * (size_t)blockIdx.x + 1] = b;
  }
}

// Basis size up to which the thick solver's expansion steps use the fused
// pass k_ritz_expand (its per-thread dot accumulators are registers).
constexpr int kFuseMax = 16;
// k_ritz_expand's |W|^2, |R|^2 partials, past its dot partials in the
// context's partial-sum buffer
constexpr size_t kPwOffset = (size_t)kRedBlocks * (kFuseMax + 1);

// One pass over the basis for an expansion step of the thick solver with
// the diagonal preconditioner: W = sum_k U[k] C_k and R = E0 W -
// sum_k U[k] HC_k as k_ritz (R is not stored), the diagonal step
// T = R/(E0 - d) with k_precond's regularization (T is the unnormalized
// new direction: Gram-Schmidt normalizes it), and per block the partial
// dots <C_k|T> (k < nv) and <T|T> at part[b*(nv+1) + k] (the layout
// k_finish_dgks reads) and |W|^2, |R|^2 at pw[2b], pw[2b+1].  Replaces
// k_ritz, k_scale2, k_precond and the first Gram-Schmidt k_dots (their
// separate reads of R, T and the basis).  nv <= kFuseMax.
__global__ void __launch_bounds__(kRedThreads) k_ritz_expand(VecList Cv, VecList HCv, int nv,
                                                             const double *__restrict__ U,
                                                             const double *__restrict__ E,
                                                             const double *__restrict__ hii, double eps_reg,
                                                             size_t n, double *__restrict__ W,
                                                             double *__restrict__ T, double *__restrict__ part,
                                                             double *__restrict__ pw) {
  constexpr int NS = kFuseMax + 3;  // dots with C_k, then |T|^2, |W|^2, |R|^2
  __shared__ double sh[kRedThreads / 64][NS];
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const double E0 = E[0];
  double acc[NS];
#pragma unroll
  for (int k = 0; k < NS; k++) acc[k] = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double cv[kFuseMax];
    double w = 0.0, r = 0.0;
    // basis loads in groups of 8 issued together, sums in index order (as
    // k_ritz); lanes past nv reload vector 0 and are not summed
#pragma unroll
    for (int k0 = 0; k0 < kFuseMax; k0 += 8) {
      if (k0 < nv) {
        double hv[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
          const int k = k0 + j < nv ? k0 + j : 0;
          cv[k0 + j] = Cv.p[k][i];
          hv[j] = vget(HCv, k, i);
        }
#pragma unroll
        for (int j = 0; j < 8; j++)
          if (k0 + j < nv) {
            w += cv[k0 + j] * U[k0 + j];
            r += hv[j] * (-U[k0 + j]);
          }
      }
    }
    r += E0 * w;
    const double den = E0 - hii[i];
    const double t = fabs(den) > eps_reg ? r / den : r / (den - eps_reg);
    W[i] = w;
    T[i] = t;
#pragma unroll
    for (int k = 0; k < kFuseMax; k++)
      if (k < nv) acc[k] += cv[k] * t;
    acc[kFuseMax] += t * t;
    acc[kFuseMax + 1] += w * w;
    acc[kFuseMax + 2] += r * r;
  }
  // fixed-order block sums: butterfly within each wavefront, then the 4
  // wavefronts in order
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NS; k++) {
    double v = acc[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) sh[wv][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < NS) {
    const int k = threadIdx.x;
    double v = sh[0][k];
    for (int q = 1; q < kRedThreads / 64; q++) v += sh[q][k];
    if (k < nv) part[(size_t)blockIdx.x * (nv + 1) + k] = v;
    else if (k == kFuseMax) part[(size_t)blockIdx.x * (nv + 1) + nv] = v;
    else if (k == kFuseMax + 1) pw[2 * (size_t)blockIdx.x] = v;
    else if (k == kFuseMax + 2) pw[2 * (size_t)blockIdx.x + 1] = v;
  }
}

// scal[0] = |R|, scal[1] = |W| from k_ritz_expand's block partials
__global__ void __launch_bounds__(64) k_finish_norms(int nblk, const double *__restrict__ pw,
                                                     double *__restrict__ scal) {
  const double nw = wave_sum(pw, nblk, 2, 0);
  const double nr = wave_sum(pw, nblk, 2, 1);
  if (threadIdx.x == 0) {
    scal[0] = sqrt(nr);
    scal[1] = sqrt(nw);
  }
}

// v /= *nrm
__global__ void __launch_bounds__(kRedThreads) k_scale_by(double *__restrict__ v, size_t n,
                                                          const double *__restrict__ nrm) {
  const double f = 1.0 / nrm[0];
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) v[i] *= f;
}

// W /= |W| ; R /= |R| ; scal[0] = |R|, scal[1] = |W|.
__global__ void __launch_bounds__(kRedThreads) k_scale2(double *__restrict__ W, double *__restrict__ R,
                                                        size_t n, const double *__restrict__ part,
                                                        int nblk, double *__restrict__ scal) {
  __shared__ double f[2];
  if (threadIdx.x < 64) {
    const double nw = sqrt(wave_sum(part, nblk, 2, 0));
    const double nr = sqrt(wave_sum(part, nblk, 2, 1));
    if (threadIdx.x == 0) {
      f[0] = 1.0 / nw;
      f[1] = 1.0 / nr;
      if (blockIdx.x == 0) {
        scal[0] = nr;
        scal[1] = nw;
      }
    }
  }
  __syncthreads();
  const double fw = f[0], fr = f[1];
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    W[i] *= fw;
    R[i] *= fr;
  }
}

// C = R/(E0-d), with E0-d-eps when |E0-d| <= eps (davidson.h:861-868).
__global__ void __launch_bounds__(kRedThreads) k_precond(const double *__restrict__ R,
                                                         const double *__restrict__ d,
                                                         const double *__restrict__ E, size_t n,
                                                         double eps_reg, double *__restrict__ C) {
  const double E0 = E[0];
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double den = E0 - d[i];
    C[i] = fabs(den) > eps_reg ? R[i] / den : R[i] / (den - eps_reg);
  }
}

// c -= (sum of partials) * x   (one modified Gram-Schmidt projection)
__global__ void __launch_bounds__(kRedThreads) k_axpy_part(double *__restrict__ c,
                                                           const double *__restrict__ x, size_t n,
                                                           const double *__restrict__ part,
                                                           int nblk) {
  __shared__ double ol;
  if (threadIdx.x < 64) {
    const double v = wave_sum(part, nblk, 1, 0);
    if (threadIdx.x == 0) ol = -v;
  }
  __syncthreads();
  const double o = ol;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    c[i] += x[i] * o;
}

// c *= 1/sqrt(sum of partials)
__global__ void __launch_bounds__(kRedThreads) k_scale_part(double *__restrict__ c, size_t n,
                                                            const double *__restrict__ part,
                                                            int nblk) {
  __shared__ double f;
  if (threadIdx.x < 64) {
    const double v = wave_sum(part, nblk, 1, 0);
    if (threadIdx.x == 0) f = 1.0 / sqrt(v);
  }
  __syncthreads();
  const double g = f;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) c[i] *= g;
}

// *out = fixed-order sum of the nblk partials (one wavefront)
__global__ void __launch_bounds__(64) k_finish_scalar(int nblk, const double *__restrict__ part, double *out) {
  const double v = wave_sum(part, nblk, 1, 0);
  if (threadIdx.x == 0) *out = v;
}

static void dots(Context &c, const double *const *X, int nv, const double *y, double *part,
                 const int *gate = nullptr) {
  VecList L{};
  for (int k = 0; k < nv; k++) L.p[k] = X[k];
  hipLaunchKernelGGL(k_dots, dim3(kRedBlocks), dim3(kRedThreads), 0, c.stream, L, nv, y, c.N, part, gate);
  check_launch("k_dots");
}
static void dots(Context &c, const VecList &L, int nv, const double *y, double *part) {
  hipLaunchKernelGGL(k_dots, dim3(kRedBlocks), dim3(kRedThreads), 0, c.stream, L, nv, y, c.N, part,
                     (const int *)nullptr);
  check_launch("k_dots");
}

double device_dot(Context &c, const double *x, const double *y, size_t n) {
  (void)n;
  const double *X[1] = {x};
  dots(c, X, 1, y, c.part);
  double *scal = c.dsmall + kScalarOffset + 8;
  hipLaunchKernelGGL(k_finish_scalar, dim3(1), dim3(64), 0, c.stream, kRedBlocks, c.part, scal);
  check_launch("k_finish_scalar");
  double r = 0.0;
  SBDX_HIP(hipMemcpyAsync(&r, scal, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  SBDX_HIP(hipStreamSynchronize(c.stream));
  return r;
}

// out[k] = fixed-order sum over blocks of part[blk*nv + k]
__global__ void __launch_bounds__(kFinishThreads) k_finish_vec(int nv, int nblk, const double *__restrict__ part,
                                                               double *__restrict__ out, const int *__restrict__ gate) {
  if (gate && *gate == 0) return;
  for (int k = threadIdx.x >> 6; k < nv; k += kFinishThreads / 64) {
    const double s = wave_sum(part, nblk, nv, k);
    if ((threadIdx.x & 63) == 0) out[k] = s;
  }
}

// t -= sum_k coef[k] X_k (one classical Gram-Schmidt pass with coefficients
// already reduced)
__global__ void __launch_bounds__(kRedThreads) k_sub_combo(VecList X, int nv, const double *__restrict__ coef,
                                                           double *__restrict__ t, size_t n,
                                                           const int *__restrict__ gate) {
  if (gate && *gate == 0) return;  // uniform per block, before any barrier
  __shared__ double cf[kNbMax];
  if (threadIdx.x < nv) cf[threadIdx.x] = coef[threadIdx.x];
  __syncthreads();
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double acc = t[i];
    for (int k0 = 0; k0 < nv; k0 += 8) {  // loads of 8 vectors issue together
      const int m = nv - k0 < 8 ? nv - k0 : 8;
      double v[8];
#pragma unroll
      for (int j = 0; j < 8; j++) v[j] = X.p[j < m ? k0 + j : k0][i];
#pragma unroll
      for (int j = 0; j < 8; j++)
        if (j < m) acc -= cf[k0 + j] * v[j];
    }
    t[i] = acc;
  }
}

// First Gram-Schmidt pass of t against nv basis vectors: part holds the
// partial dots of the nv vectors and, as entry nv, of t with itself.
// coef[k] = <X_k|t>; *gate = 1 when the pass leaves less than half of
// |t|^2 (|t|^2 - sum coef^2 < |t|^2 / 2), so that a second pass is needed
// to keep t orthogonal (Daniel, Gragg, Kaufman and Stewart 1976,
// threshold 1/sqrt(2) on the norm), else 0.  One block of kFinishThreads.
__global__ void __launch_bounds__(kFinishThreads) k_finish_dgks(int nv, int nblk, const double *__restrict__ part,
                                                                double *__restrict__ coef, int *__restrict__ gate) {
  __shared__ double s[kNbMax + 1];
  for (int k = threadIdx.x >> 6; k <= nv; k += kFinishThreads / 64) {
    const double a = wave_sum(part, nblk, nv + 1, k);
    if ((threadIdx.x & 63) == 0) {
      s[k] = a;
      if (k < nv) coef[k] = a;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double hh = 0.0;
    for (int j = 0; j < nv; j++) hh += s[j] * s[j];
    *gate = (s[nv] - hh < 0.5 * s[nv]) ? 1 : 0;
  }
}

// out_j = sum_k U[k + ld*(j0+j)] in_k for j < nout (Ritz vectors and their
// H-products from the subspace basis)
__global__ void __launch_bounds__(kRedThreads) k_combine(VecList in, int nin, const double *__restrict__ U,
                                                         int ld, int j0, int nout, OutList out, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double acc[8];
#pragma unroll
    for (int j = 0; j < 8; j++) acc[j] = 0.0;
    for (int k0 = 0; k0 < nin; k0 += 8) {  // loads of 8 inputs issue together
      const int m = nin - k0 < 8 ? nin - k0 : 8;
      double v[8];
#pragma unroll
      for (int q = 0; q < 8; q++) v[q] = vget(in, q < m ? k0 + q : k0, i);
#pragma unroll
      for (int q = 0; q < 8; q++)
        if (q < m) {
#pragma unroll
          for (int j = 0; j < 8; j++)
            if (j < nout) acc[j] += U[k0 + q + (size_t)ld * (j0 + j)] * v[q];
        }
    }
#pragma unroll
    for (int j = 0; j < 8; j++)
      if (j < nout) out.p[j][i] = acc[j];
  }
}

// R = E*W - HW with E = <W|HW> (from partials); partial |R|^2 per block.
__global__ void __launch_bounds__(kRedThreads) k_exact_resid(const double *__restrict__ W,
                                                             const double *__restrict__ HW, size_t n,
                                                             const double *__restrict__ epart, int nblk,
                                                             double *__restrict__ R,
                                                             double *__restrict__ part,
                                                             double *__restrict__ Eout) {
  __shared__ double sh[kRedThreads];
  __shared__ double Es;
  if (threadIdx.x < 64) {
    const double v = wave_sum(epart, nblk, 1, 0);
    if (threadIdx.x == 0) {
      Es = v;
      if (blockIdx.x == 0) *Eout = v;
    }
  }
  __syncthreads();
  const double E = Es;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double r = E * W[i] - HW[i];  // same sign as k_ritz's residual
    R[i] = r;
    s += r * r;
  }
  const double t = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

static void ensure_subspace(Context &c, int nb) {
  while ((int)c.C.size() < nb) {
    c.C.push_back((double *)dalloc(c, sizeof(double) * c.N));
    c.HC.push_back((double *)dalloc(c, sizeof(double) * c.N));
  }
  if (!c.R) c.R = (double *)dalloc(c, sizeof(double) * c.N);
}

int davidson_solve(Context &c, int max_it, int nb, double eps, double max_time, double *energy,
                   int *n_steps) {
  if (nb < 1 || nb > kNbMax) throw Error(1, "davidson: nb must be in [1, 64]");
  ensure_subspace(c, nb);
  const double eps_reg = 1.0e-12;
  const int ld = kNbMax;
  double *H = c.dsmall, *U = c.dsmall + ld * ld, *E = c.dsmall + 2 * ld * ld;
  double *scal = c.dsmall + kScalarOffset;
  const size_t N = c.N;
  const dim3 g(kRedBlocks), b(kRedThreads);
  c.trace.clear();
  auto t_start = std::chrono::steady_clock::now();
  std::vector<double> step_times;
  int steps = 0;
  bool done = false, stopped = false;
  double E0 = 0.0;
  for (int it = 0; it < max_it && !done && !stopped; it++) {
    SBDX_HIP(hipMemcpyAsync(c.C[0], c.W, sizeof(double) * N, hipMemcpyDeviceToDevice, c.stream));
    for (int ib = 0; ib < nb; ib++) {
      auto s0 = std::chrono::steady_clock::now();
      // Mixed precision: the restart vector (ib == 0) gets the exact FP64
      // product, so every cycle starts from the exact FP64 residual of W;
      // the correction vectors use the FP32 product, whose error enters W
      // only through coefficients that shrink with the residual.
      const bool approx = c.precision == 1 && ib > 0;
      if (approx)
        sigma_alpha_first_f32(c, c.C[ib], c.HC[ib]);
      else
        sigma_alpha_first(c, c.C[ib], c.HC[ib]);
      {
        ProfScope ps(c, "dav_subspace");
        // H[k][ib] = <C_k|H|C_ib>.  With an FP32 product in HC[ib] the
        // coupling to the restart vector is taken as C_ib . HC[k] instead:
        // HC[0] is exact, and <W|H|t> = t . (HW) is of the size of the
        // residual, which the FP32 error of H t (~1e-7 |H|) would swamp.
        if (approx)
          dots(c, c.HC.data(), ib + 1, c.C[ib], c.part);
        else
          dots(c, c.C.data(), ib + 1, c.HC[ib], c.part);
        hipLaunchKernelGGL(k_finish_hcol, dim3(1), dim3(kFinishThreads), 0, c.stream, ib + 1, kRedBlocks,
                           c.part, H, ld, ib);
        check_launch("k_finish_hcol");
        subspace_eig(c, ib + 1, ld, H, U, E);
      }
      {
        ProfScope ps(c, "dav_ritz");
        VecList Cv{}, HCv{};
        for (int k = 0; k <= ib; k++) {
          Cv.p[k] = c.C[k];
          HCv.p[k] = c.HC[k];
        }
        hipLaunchKernelGGL(k_ritz, g, b, 0, c.stream, Cv, HCv, ib + 1, U, E, N, c.W, c.R, c.part);
        check_launch("k_ritz");
        hipLaunchKernelGGL(k_scale2, g, b, 0, c.stream, c.W, c.R, N, c.part, kRedBlocks, scal);
        check_launch("k_scale2");
      }
      double hs[2];
      SBDX_HIP(hipMemcpyAsync(&hs[0], E, sizeof(double), hipMemcpyDeviceToHost, c.stream));
      SBDX_HIP(hipMemcpyAsync(&hs[1], scal, sizeof(double), hipMemcpyDeviceToHost, c.stream));
      SBDX_HIP(hipStreamSynchronize(c.stream));
      E0 = hs[0];
      const double nr = hs[1];
      c.trace.push_back(E0);
      c.trace.push_back(nr);
      steps++;
      if (c.verbose) {
        fprintf(stderr, "[sbdx] davidson %d.%d E0 = %.15f |r| = %.3e\n", it, ib, E0, nr);
        fflush(stderr);
      }
      if (nr < eps) {
        if (approx) break;  // confirm on the exact residual: restart from W
        done = true;
        break;
      }
      if (ib < nb - 1) {
        ProfScope ps(c, "dav_orth");
        double *cn = c.C[ib + 1];
        hipLaunchKernelGGL(k_precond, g, b, 0, c.stream, c.R, c.hii, E, N, eps_reg, cn);
        check_launch("k_precond");
        for (int kb = 0; kb <= ib; kb++) {
          const double *X[1] = {c.C[kb]};
          dots(c, X, 1, cn, c.part);
          hipLaunchKernelGGL(k_axpy_part, g, b, 0, c.stream, cn, c.C[kb], N, c.part, kRedBlocks);
          check_launch("k_axpy_part");
        }
        const double *X[1] = {cn};
        dots(c, X, 1, cn, c.part);
        hipLaunchKernelGGL(k_scale_part, g, b, 0, c.stream, cn, N, c.part, kRedBlocks);
        check_launch("k_scale_part");
      }
      auto s1 = std::chrono::steady_clock::now();
      step_times.push_back(std::chrono::duration<double>(s1 - s0).count());
      double avg = 0.0;
      for (double t : step_times) avg += t;
      avg /= (double)step_times.size();
      const double elapsed = std::chrono::duration<double>(s1 - t_start).count();
      if (elapsed + avg > max_time) {
        stopped = true;
        break;
      }
    }
  }
  SBDX_HIP(hipStreamSynchronize(c.stream));
  if (energy) *energy = E0;
  if (n_steps) *n_steps = steps;
  return done ? 0 : 5;
}

// Thick-restart Davidson (no counterpart in the reference, which restarts from
// W alone).  Per step: one new basis vector from the preconditioned residual
// (same preconditioner, MGS and normalization as above), its product, the new
// subspace column, Rayleigh-Ritz.  When the basis is full, or when the
// in-cycle residual estimate falls below eps, the solver restarts:
//   1. exact FP64 product of the Ritz vector W, E = <W|HW>, R = HW - E W:
//      convergence is decided only on this exact |R|;
//   2. the basis keeps W (with its exact product) and the next keep-1 Ritz
//      vectors, whose products are linear combinations of the old ones.
// With mixed != 0 the in-cycle products are FP32 (sigma_alpha_first_f32);
// their error enters W only through coefficients that shrink with the
// residual, and every restart re-anchors the basis on an exact FP64 product.
int davidson_solve_thick(Context &c, int max_steps, int nb, int keep, double eps, double max_time,
                         bool mixed, double *energy, int *n_steps) {
  if (nb < 3 || nb > kNbMax) throw Error(1, "davidson: nb must be in [3, 64]");
  const int m = std::max(1, std::min(keep, std::min(8, nb - 2)));
  ensure_subspace(c, nb);
  while ((int)c.spareY.size() < m) {
    c.spareY.push_back((double *)dalloc(c, sizeof(double) * c.N));
    c.spareHY.push_back((double *)dalloc(c, sizeof(double) * c.N));
  }
  const double eps_reg = 1.0e-12;
  const int ld = kNbMax;
  double *H = c.dsmall, *U = c.dsmall + ld * ld, *E = c.dsmall + 2 * ld * ld;
  double *scal = c.dsmall + kScalarOffset;
  const size_t N = c.N;
  const dim3 g(kRedBlocks), b(kRedThreads);
  c.trace.clear();
  auto t_start = std::chrono::steady_clock::now();
  auto elapsed = [&]() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  };
  auto host2 = [&](const double *d0, const double *d1, double &a, double &bb) {
    SBDX_HIP(hipMemcpyAsync(&a, d0, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    SBDX_HIP(hipMemcpyAsync(&bb, d1, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    SBDX_HIP(hipStreamSynchronize(c.stream));
  };
  // Column col of the subspace matrix as C[col] . HC[k], k <= col: the
  // older vector's product is used, so the coupling of every later vector to
  // the restart vector (whose product is exact FP64) stays exact in mixed
  // mode; an FP32 error of ~1e-7 |H| there would exceed the residual-sized
  // coupling <W|H|t> = t . r near convergence.
  // Products of FP32-product columns are kept in FP32 (c.HCf) in the
  // symmetric mixed solve: bit k of hc32 marks column k.  The values are
  // those an FP64 copy would hold (an FP32 result widened), at half the
  // bytes for every later pass over the products.
  uint64_t hc32 = 0;
  auto hc_list = [&](int nv) {
    VecList L{};
    for (int k = 0; k < nv; k++) {
      L.p[k] = c.HC[k];
      L.pf[k] = k < (int)c.HCf.size() ? c.HCf[k] : nullptr;
    }
    L.f32 = nv >= 64 ? hc32 : (hc32 & ((1ull << nv) - 1ull));
    return L;
  };
  auto add_column = [&](int col) {
    dots(c, hc_list(col + 1), col + 1, c.C[col], c.part);
    hipLaunchKernelGGL(k_finish_hcol, dim3(1), dim3(kFinishThreads), 0, c.stream, col + 1, kRedBlocks, c.part,
                       H, ld, col);
    check_launch("k_finish_hcol");
  };
  // spin-flip symmetric products when asked for: the start vector must lie
  // in the sector c.sym_par (x = x^T even, x = -x^T odd, in the alpha-first
  // layout); the diagonal step keeps every later vector there (hii is
  // symmetric), the learned step is projected back
  const bool sym = c.symmetric != 0;
  if (sym) {
    if (!sigma_sym_available(c))
      throw Error(1, "symmetric mode needs equal alpha and beta string sets and room for the dense blocks");
    const double defect = sym_defect(c, c.W, c.sym_par);
    if (defect > 1e-6)
    {
      char msg[160];
      snprintf(msg, sizeof msg, "symmetric mode: the start vector is not in the spin-flip %s sector (defect %.3e)",
               c.sym_par < 0 ? "odd" : "even", defect);
      throw Error(1, msg);
    }
    symmetrize(c, c.W, c.sym_par);  // rounding-level defect of a vector from elsewhere
  }
  auto apply = [&](const double *x, double *y, bool approx) {
    if (sym && approx)
      sigma_sym_f32(c, x, y);
    else if (sym)
      sigma_sym(c, x, y);
    else if (approx)
      sigma_alpha_first_f32(c, x, y);
    else
      sigma_alpha_first(c, x, y);
  };
  // coarse start: the start vector restricted to the first coarse_m strings
  // of each set is solved on that sub-problem (same solver settings,
  // diagonal preconditioner) to coarse_eps and lifted back (zero elsewhere;
  // the basis sign depends only on the two strings, so no re-signing)
  c.coarse_products = 0;
  c.coarse_energy = 0.0;
  c.coarse_seconds = 0.0;
  const size_t mc = c.coarse_m;
  if (mc > 0 && mc < c.na && mc < c.nb) {
    ProfScope ps(c, "coarse_start");
    const auto tc = std::chrono::steady_clock::now();
    // the sub-problem's context (tables, dense blocks, tuned kernels) is kept
    // for the next solve of the same size, e.g. the other spin-flip sector
    if (c.coarse_sub && c.coarse_sub_m != mc) {
      destroy_context(c.coarse_sub);
      c.coarse_sub = nullptr;
    }
    if (!c.coarse_sub) {
      c.coarse_sub = make_sub_context(c, mc);
      c.coarse_sub_m = mc;
    }
    Context *sub = c.coarse_sub;
    sub->mixed_form = c.mixed_form;
    sub->precision = c.precision;
    sub->restart_keep = c.restart_keep;
    // nested levels: the sub-problem starts from a quarter of its strings
    sub->coarse_m = mc / 4 >= 256 ? mc / 4 : 0;
    sub->coarse_eps = c.coarse_eps;
    sub->symmetric = (sym && sigma_sym_available(*sub)) ? (c.sym_par < 0 ? 2 : 1) : 0;
    sub->sym_par = c.sym_par;
    // copies between the two contexts run on the consumer's stream after the
    // producer's stream has drained (both are non-blocking streams)
    SBDX_HIP(hipStreamSynchronize(c.stream));
    SBDX_HIP(hipMemcpy2DAsync(sub->W, sizeof(double) * mc, c.W, sizeof(double) * c.nb, sizeof(double) * mc, mc,
                              hipMemcpyDeviceToDevice, sub->stream));
    const double n2 = device_dot(*sub, sub->W, sub->W, sub->N);
    if (n2 > 1e-20) {
      double es = 0.0;
      int ns = 0;
      davidson_solve_thick(*sub, max_steps, nb, keep, c.coarse_eps, max_time, mixed, &es, &ns);
      SBDX_HIP(hipStreamSynchronize(sub->stream));
      SBDX_HIP(hipMemsetAsync(c.W, 0, sizeof(double) * N, c.stream));
      SBDX_HIP(hipMemcpy2DAsync(c.W, sizeof(double) * c.nb, sub->W, sizeof(double) * mc, sizeof(double) * mc,
                                mc, hipMemcpyDeviceToDevice, c.stream));
      SBDX_HIP(hipStreamSynchronize(c.stream));
      c.coarse_products = ns;
      c.coarse_energy = es;
    }
    c.coarse_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - tc).count();
    if (c.verbose) {
      fprintf(stderr, "[sbdx] coarse start on %zu x %zu strings: E = %.12f, %d products, %.3f s\n", mc, mc,
              c.coarse_energy, c.coarse_products, c.coarse_seconds);
      fflush(stderr);
    }
  }
  // start: C[0] = W / |W|, exact product
  SBDX_HIP(hipMemcpyAsync(c.C[0], c.W, sizeof(double) * N, hipMemcpyDeviceToDevice, c.stream));
  {
    const double *X[1] = {c.C[0]};
    dots(c, X, 1, c.C[0], c.part);
    hipLaunchKernelGGL(k_scale_part, g, b, 0, c.stream, c.C[0], N, c.part, kRedBlocks);
    check_launch("k_scale_part");
  }
  // Anchor products (the start and the restarts) are exact so that the
  // subspace couples every correction to an exact product once the
  // residual nears the FP32 floor.  They stay FP32 (errors ~1e-7 |H|)
  //  - while the residual r of the step lies ten times above that floor
  //    (kFp32Floor |E0|): the corrections' own FP32 products dominate the
  //    subspace's error there, and a later restart anchors exactly;
  //  - in an abandonable stage (abandon_above set: another spin-flip
  //    sector's result lies at that level) while E0 lies more than
  //    kAnchorMargin |E0| above the level: the stage is only needed to show
  //    that its lowest state lies above, which FP32 values show as well.
  // Convergence is only declared on an exact anchor.
  auto approx_anchor = [&](double e, double r) {
    if (!mixed) return false;
    const double ae = std::max(1.0, std::fabs(e));
    return r > 10.0 * kFp32Floor * ae || (c.abandon_above < 1e299 && e - c.abandon_above > kAnchorMargin * ae);
  };
  // the lifted coarse vector's Rayleigh quotient is the coarse energy (the
  // sub-problem is H restricted to the coarse block); its residual is not
  // known before this product, so only the second rule applies
  apply(c.C[0], c.HC[0], c.coarse_products > 0 && approx_anchor(c.coarse_energy, 0.0));
  hc32 = 0;
  add_column(0);
  int d = 0, steps = 1;
  bool done = false, fresh = true, abandoned = false;
  double E0 = 0.0, nr = 0.0;
  std::vector<double> rhist;
  double prev_exact = -1.0;  // exact residual of the previous restart
  c.fp64_switch_step = -1;
  int diag_left = 0;
  c.mlp.used_mlp = c.mlp.used_diag = 0;
  c.mlp.rec_n = 0;
  c.mlp.rec_E.clear();
  while (true) {
    {
      ProfScope ps(c, "dav_subspace");
      subspace_eig(c, d + 1, ld, H, U, E);
    }
    // An expansion step with the diagonal step takes the new direction and
    // its first Gram-Schmidt dots from the Ritz pass itself (k_ritz_expand):
    // no restart is due at this basis size, no training pairs are being
    // recorded (they need R) and the network is not the preconditioner.
    // W then stays unnormalized (unit up to rounding) until the step ends
    // other than by expansion, when normalize_w scales it exactly.
    const bool fuse = d + 1 <= kFuseMax && d < nb - 1 && c.mlp.record_cap == 0 &&
                      !(c.precond == 1 && c.mlp.valid);
    bool have_cn = false;
    {
      ProfScope ps(c, "dav_ritz");
      VecList Cv{};
      for (int k = 0; k <= d; k++) Cv.p[k] = c.C[k];
      const VecList HCv = hc_list(d + 1);
      if (fuse) {
        hipLaunchKernelGGL(k_ritz_expand, g, b, 0, c.stream, Cv, HCv, d + 1, U, E, c.hii, eps_reg, N, c.W,
                           c.C[d + 1], c.part, c.part + kPwOffset);
        check_launch("k_ritz_expand");
        hipLaunchKernelGGL(k_finish_norms, dim3(1), dim3(64), 0, c.stream, kRedBlocks, c.part + kPwOffset, scal);
        check_launch("k_finish_norms");
      } else {
        hipLaunchKernelGGL(k_ritz, g, b, 0, c.stream, Cv, HCv, d + 1, U, E, N, c.W, c.R, c.part);
        check_launch("k_ritz");
        hipLaunchKernelGGL(k_scale2, g, b, 0, c.stream, c.W, c.R, N, c.part, kRedBlocks, scal);
        check_launch("k_scale2");
      }
    }
    auto normalize_w = [&]() {
      if (!fuse) return;
      hipLaunchKernelGGL(k_scale_by, g, b, 0, c.stream, c.W, N, scal + 1);
      check_launch("k_scale_by");
    };
    host2(E, scal, E0, nr);
    c.trace.push_back(E0);
    c.trace.push_back(nr);
    if (c.verbose) {
      fprintf(stderr, "[sbdx] thick step %d dim %d E0 = %.15f |r| ~ %.3e\n", steps, d + 1, E0, nr);
      fflush(stderr);
    }
    const bool exact_now = !mixed;  // every product in the basis is FP64
    if (nr < eps && exact_now) {
      normalize_w();
      done = true;
      break;
    }
    // With the unit Ritz vector u = sum c_i v_i over this sector's
    // eigenvectors, |r|^2 >= |c_1|^2 (E0 - lambda_1)^2, so E0 - lambda_1 <=
    // |r| / |c_1|.  Were the sector's lowest eigenvalue below abandon_above,
    // E0 - abandon_above > 10 |r| would need |c_1| < 1/10: while u keeps an
    // overlap of at least 1/10 with the lowest state (the state the
    // iteration converges to), the test shows that this sector's lowest
    // state lies above the level and the caller has no use for it
    if (E0 - c.abandon_above > 10.0 * nr) {
      normalize_w();
      abandoned = true;
      break;
    }
    const bool timeout = elapsed() > max_time || steps >= max_steps;
    if (d == nb - 1 || (nr < eps && !fresh) || timeout) {
      // ---- restart with an exact residual check
      ProfScope ps(c, "dav_restart");
      normalize_w();
      double *HW = c.spareHY[0];
      const bool approx = approx_anchor(E0, nr);
      apply(c.W, HW, approx);
      steps++;
      const double *X[1] = {c.W};
      dots(c, X, 1, HW, c.part);
      double *Eex = scal + 2;
      hipLaunchKernelGGL(k_exact_resid, g, b, 0, c.stream, c.W, HW, N, c.part, kRedBlocks, c.R,
                         c.part + kRedBlocks, Eex);
      check_launch("k_exact_resid");
      hipLaunchKernelGGL(k_finish_scalar, dim3(1), dim3(64), 0, c.stream, kRedBlocks,
                         c.part + kRedBlocks, scal + 3);
      check_launch("k_finish_scalar");
      double e_ex, r2;
      host2(Eex, scal + 3, e_ex, r2);
      E0 = e_ex;
      nr = std::sqrt(r2);
      c.trace.push_back(E0);
      c.trace.push_back(nr);
      if (c.verbose) {
        fprintf(stderr, "[sbdx] restart step %d E0 = %.15f exact |r| = %.3e\n", steps, E0, nr);
        fflush(stderr);
      }
      if (nr < eps && !approx) {
        done = true;
        break;
      }
      if (timeout) break;
      // precision switch: FP32 product errors (~1e-7 |H| per product) bound
      // how far FP32 corrections can take the residual, and below that
      // floor they feed errors back through the kept Ritz vectors.  When an
      // exact restart residual fails to halve while already small against
      // the FP32 noise level of this spectrum, the rest of the solve uses
      // FP64 products, and this restart keeps W alone (its product exact).
      int keep_now = m;
      if (mixed && !approx && prev_exact > 0.0 && nr > 0.5 * prev_exact && nr < kFp32Floor * std::fabs(E0)) {
        mixed = false;
        keep_now = 1;
        c.fp64_switch_step = steps;
        if (c.verbose) {
          fprintf(stderr, "[sbdx] FP32 product floor reached at |r| = %.3e: FP64 products from step %d\n", nr,
                  steps);
          fflush(stderr);
        }
      }
      prev_exact = approx ? -1.0 : nr;  // an FP32 anchor residual is no baseline
      // R /= |R| for the preconditioner (the reference normalizes R too)
      {
        const double *Y[1] = {c.R};
        dots(c, Y, 1, c.R, c.part);
        hipLaunchKernelGGL(k_scale_part, g, b, 0, c.stream, c.R, N, c.part, kRedBlocks);
        check_launch("k_scale_part");
      }
      // thick restart basis: W (exact product) + Ritz vectors 1..keep_now-1
      if (keep_now > 1) {
        VecList Cv{};
        for (int k = 0; k <= d; k++) Cv.p[k] = c.C[k];
        const VecList HCv = hc_list(d + 1);
        OutList Yo{}, HYo{};
        for (int j = 1; j < keep_now; j++) {
          Yo.p[j - 1] = c.spareY[j];
          HYo.p[j - 1] = c.spareHY[j];
        }
        hipLaunchKernelGGL(k_combine, g, b, 0, c.stream, Cv, d + 1, U, ld, 1, keep_now - 1, Yo, N);
        check_launch("k_combine");
        hipLaunchKernelGGL(k_combine, g, b, 0, c.stream, HCv, d + 1, U, ld, 1, keep_now - 1, HYo, N);
        check_launch("k_combine");
      }
      SBDX_HIP(hipMemcpyAsync(c.C[0], c.W, sizeof(double) * N, hipMemcpyDeviceToDevice, c.stream));
      std::swap(c.HC[0], c.spareHY[0]);
      for (int j = 1; j < keep_now; j++) {
        std::swap(c.C[j], c.spareY[j]);
        std::swap(c.HC[j], c.spareHY[j]);
      }
      d = keep_now - 1;
      hc32 &= ~((1ull << keep_now) - 1ull);  // the kept columns' products are FP64
      for (int j = 0; j <= d; j++) add_column(j);
      fresh = true;
      // expand from the exact residual (E0 = exact energy for the preconditioner)
      SBDX_HIP(hipMemcpyAsync(E, Eex, sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    } else {
      have_cn = fuse;  // the Ritz pass wrote the direction and its dots
    }
    // ---- expand: preconditioned residual, MGS against the basis, normalize
    rhist.push_back(nr);
    if (c.mlp.record_cap > 0 && c.mlp.rec_n < c.mlp.record_cap) {
      // training pairs: residual, Ritz vector and Ritz value of this step
      const int k = c.mlp.rec_n++;
      if ((int)c.mlp.rec_R.size() <= k) {
        c.mlp.rec_R.push_back((double *)dalloc(c, sizeof(double) * N));
        c.mlp.rec_W.push_back((double *)dalloc(c, sizeof(double) * N));
      }
      SBDX_HIP(hipMemcpyAsync(c.mlp.rec_R[k], c.R, sizeof(double) * N, hipMemcpyDeviceToDevice, c.stream));
      SBDX_HIP(hipMemcpyAsync(c.mlp.rec_W[k], c.W, sizeof(double) * N, hipMemcpyDeviceToDevice, c.stream));
      c.mlp.rec_E.push_back(E0);
    }
    // learned step unless the residual stalls (no 10% drop over 3 steps):
    // then the diagonal step for the next 3 expansions
    const size_t nh = rhist.size();
    if (nh >= 4 && rhist[nh - 1] > 0.9 * rhist[nh - 4] && diag_left == 0) diag_left = 3;
    const bool use_mlp = c.precond == 1 && c.mlp.valid && diag_left == 0;
    if (diag_left > 0) diag_left--;
    {
      ProfScope ps(c, use_mlp ? "precond_mlp" : "dav_orth");
      double *cn = c.C[d + 1];
      if (use_mlp) {
        mlp_precond(c, c.R, E, cn);
        if (sym) symmetrize(c, cn, c.sym_par);
        c.mlp.used_mlp++;
      } else {
        if (!have_cn) {
          hipLaunchKernelGGL(k_precond, g, b, 0, c.stream, c.R, c.hii, E, N, eps_reg, cn);
          check_launch("k_precond");
        }
        c.mlp.used_diag++;
      }
    }
    {
      // classical Gram-Schmidt against the d+1 basis vectors (one multi-dot
      // and one combined update per pass); the second pass runs only when
      // the first removed more than half of |cn|^2, decided on the device
      // (k_finish_dgks sets the gate the second pass's kernels read)
      ProfScope ps(c, "dav_orth");
      double *cn = c.C[d + 1];
      double *coef = c.dsmall + kCoefOffset;
      int *gate = reinterpret_cast<int *>(c.dsmall + kScalarOffset + 24);
      VecList Cv{};
      for (int k = 0; k <= d; k++) Cv.p[k] = c.C[k];
      const double *Xw[kNbMax];
      for (int k = 0; k <= d; k++) Xw[k] = c.C[k];
      Xw[d + 1] = cn;  // d + 2 <= nb <= kNbMax
      if (!have_cn) dots(c, Xw, d + 2, cn, c.part);  // else k_ritz_expand wrote them
      hipLaunchKernelGGL(k_finish_dgks, dim3(1), dim3(kFinishThreads), 0, c.stream, d + 1, kRedBlocks, c.part, coef, gate);
      check_launch("k_finish_dgks");
      hipLaunchKernelGGL(k_sub_combo, g, b, 0, c.stream, Cv, d + 1, coef, cn, N, (const int *)nullptr);
      check_launch("k_sub_combo");
      dots(c, c.C.data(), d + 1, cn, c.part, gate);
      hipLaunchKernelGGL(k_finish_vec, dim3(1), dim3(kFinishThreads), 0, c.stream, d + 1, kRedBlocks, c.part, coef,
                         (const int *)gate);
      check_launch("k_finish_vec");
      hipLaunchKernelGGL(k_sub_combo, g, b, 0, c.stream, Cv, d + 1, coef, cn, N, (const int *)gate);
      check_launch("k_sub_combo");
      const double *X[1] = {cn};
      dots(c, X, 1, cn, c.part);
      hipLaunchKernelGGL(k_scale_part, g, b, 0, c.stream, cn, N, c.part, kRedBlocks);
      check_launch("k_scale_part");
    }
    if (mixed && sym) {
      while ((int)c.HCf.size() <= d + 1) c.HCf.push_back(nullptr);
      if (!c.HCf[d + 1]) c.HCf[d + 1] = (float *)dalloc(c, sizeof(float) * N);
      sigma_sym_f32_to(c, c.C[d + 1], c.HCf[d + 1]);
      hc32 |= 1ull << (d + 1);
    } else {
      apply(c.C[d + 1], c.HC[d + 1], mixed);
      hc32 &= ~(1ull << (d + 1));
    }
    steps++;
    d++;
    add_column(d);
    fresh = false;
  }
  SBDX_HIP(hipStreamSynchronize(c.stream));
  if (energy) *energy = E0;
  if (n_steps) *n_steps = steps;
  return done ? 0 : abandoned ? 6 : 5;
}

// ------------------------------------------------------- spin-flip sectors
// Smallest hii[a*n + b] over a < b < lim (block partials: value, flat index).
__global__ void __launch_bounds__(kRedThreads) k_argmin_offdiag(size_t n, size_t lim, const double *__restrict__ hii,
                                                                double *__restrict__ part_v,
                                                                unsigned long long *__restrict__ part_i) {
  __shared__ double sv[kRedThreads];
  __shared__ unsigned long long si[kRedThreads];
  double best = 1e300;
  unsigned long long bi = ~0ull;
  const size_t N = n * n;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x) {
    if (i / n < i % n && i % n < lim && hii[i] < best) {
      best = hii[i];
      bi = i;
    }
  }
  sv[threadIdx.x] = best;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int st = kRedThreads / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      const double v = sv[threadIdx.x + st];
      const unsigned long long j = si[threadIdx.x + st];
      // ties go to the lower index, so the choice does not depend on the grid
      if (v < sv[threadIdx.x] || (v == sv[threadIdx.x] && j < si[threadIdx.x])) {
        sv[threadIdx.x] = v;
        si[threadIdx.x] = j;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part_v[blockIdx.x] = sv[0];
    part_i[blockIdx.x] = si[0];
  }
}

// The thick solve in the spin-flip sector(s) c.symmetric asks for.  Mode 3
// runs the even sector from the even part of the start vector, then the odd
// sector from its odd part (or, when that part is negligible -- the
// Hartree-Fock vector has none -- from the antisymmetric pair on the lowest
// off-diagonal diagonal element, (e_ab - e_ba)/sqrt 2, taken among the first
// coarse_m strings when the coarse start is on), and keeps the lower
// of the two.  The odd stage stops once its Ritz value lies above the even
// result by more than ten times its residual norm (near convergence the Ritz
// value exceeds its eigenvalue by ~|r|^2 / gap), so when the even sector
// holds the ground state the odd stage costs only the steps needed to show
// that.
int davidson_solve_sectors(Context &c, int max_steps, int nb, int keep, double eps, double max_time,
                           bool mixed, double *energy, int *n_steps) {
  c.sector_E[0] = c.sector_E[1] = 0.0;
  c.sector_products[0] = c.sector_products[1] = 0;
  c.sector_state[0] = c.sector_state[1] = 0;
  c.sector_chosen = 0;
  c.abandon_above = 1e300;
  if (c.symmetric != 3) {
    c.sym_par = c.symmetric == 2 ? -1 : 1;
    const int rc = davidson_solve_thick(c, max_steps, nb, keep, eps, max_time, mixed, energy, n_steps);
    if (c.symmetric != 0) {
      const int k = c.symmetric == 2 ? 1 : 0;
      c.sector_E[k] = energy ? *energy : 0.0;
      c.sector_products[k] = n_steps ? *n_steps : 0;
      c.sector_state[k] = rc == 0 ? 1 : 2;
      c.sector_chosen = c.symmetric;
    }
    return rc;
  }
  if (!sigma_sym_available(c))
    throw Error(1, "symmetric mode needs equal alpha and beta string sets and room for the dense blocks");
  const auto t0 = std::chrono::steady_clock::now();
  const size_t N = c.N, n = c.na;
  TmpBuf start(sizeof(double) * N), even(sizeof(double) * N);
  double *W0 = (double *)start.p, *We = (double *)even.p;
  SBDX_HIP(hipMemcpyAsync(W0, c.W, sizeof(double) * N, hipMemcpyDeviceToDevice, c.stream));
  const double n0 = device_dot(c, W0, W0, N);
  if (!(n0 > 0.0)) throw Error(1, "two-sector solve: zero start vector");
  // ---- even sector
  double Ee = 0.0;
  int se = 0, rce = 5;
  c.sym_par = 1;
  symmetrize(c, c.W, 1);
  const double ne = device_dot(c, c.W, c.W, N);
  if (ne > 1e-12 * n0) {
    rce = davidson_solve_thick(c, max_steps, nb, keep, eps, max_time, mixed, &Ee, &se);
    SBDX_HIP(hipMemcpyAsync(We, c.W, sizeof(double) * N, hipMemcpyDeviceToDevice, c.stream));
    c.sector_E[0] = Ee;
    c.sector_products[0] = se;
    c.sector_state[0] = rce == 0 ? 1 : 2;
  }
  // ---- odd sector
  SBDX_HIP(hipMemcpyAsync(c.W, W0, sizeof(double) * N, hipMemcpyDeviceToDevice, c.stream));
  symmetrize(c, c.W, -1);
  const double no = device_dot(c, c.W, c.W, N);
  if (!(no > 1e-6 * n0)) {
    if (n < 2) {
      // one string: the odd sector is empty
      SBDX_HIP(hipMemcpyAsync(c.W, We, sizeof(double) * N, hipMemcpyDeviceToDevice, c.stream));
      SBDX_HIP(hipStreamSynchronize(c.stream));
      c.sector_chosen = 1;
      if (energy) *energy = Ee;
      if (n_steps) *n_steps = se;
      return rce;
    }
    // within the coarse-start block when there is one, so the odd stage's
    // coarse solve starts from it too
    const size_t lim = (c.coarse_m >= 2 && c.coarse_m < n) ? c.coarse_m : n;
    TmpBuf pv(sizeof(double) * kRedBlocks), pi(sizeof(unsigned long long) * kRedBlocks);
    hipLaunchKernelGGL(k_argmin_offdiag, dim3(kRedBlocks), dim3(kRedThreads), 0, c.stream, n, lim, c.hii,
                       (double *)pv.p, (unsigned long long *)pi.p);
    check_launch("k_argmin_offdiag");
    std::vector<double> hv(kRedBlocks);
    std::vector<unsigned long long> hi(kRedBlocks);
    SBDX_HIP(hipMemcpyAsync(hv.data(), pv.p, sizeof(double) * kRedBlocks, hipMemcpyDeviceToHost, c.stream));
    SBDX_HIP(hipMemcpyAsync(hi.data(), pi.p, sizeof(unsigned long long) * kRedBlocks, hipMemcpyDeviceToHost,
                            c.stream));
    SBDX_HIP(hipStreamSynchronize(c.stream));
    double bv = 1e300;
    unsigned long long bi = ~0ull;
    for (int b = 0; b < kRedBlocks; b++)
      if (hv[b] < bv || (hv[b] == bv && hi[b] < bi)) {
        bv = hv[b];
        bi = hi[b];
      }
    const size_t a = bi / n, b = bi % n;
    const double v[2] = {std::sqrt(0.5), -std::sqrt(0.5)};
    SBDX_HIP(hipMemsetAsync(c.W, 0, sizeof(double) * N, c.stream));
    SBDX_HIP(hipMemcpyAsync(c.W + a * n + b, &v[0], sizeof(double), hipMemcpyHostToDevice, c.stream));
    SBDX_HIP(hipMemcpyAsync(c.W + b * n + a, &v[1], sizeof(double), hipMemcpyHostToDevice, c.stream));
    SBDX_HIP(hipStreamSynchronize(c.stream));
  }
  const double left =
      max_time - std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  c.sym_par = -1;
  c.abandon_above = c.sector_state[0] ? Ee : 1e300;
  double Eo = 0.0;
  int so = 0, rco = 5;
  try {
    rco = davidson_solve_thick(c, max_steps, nb, keep, eps, std::max(left, 0.0), mixed, &Eo, &so);
  } catch (...) {
    c.abandon_above = 1e300;
    c.sym_par = 1;
    throw;
  }
  c.abandon_above = 1e300;
  c.sym_par = 1;
  c.sector_E[1] = Eo;
  c.sector_products[1] = so;
  c.sector_state[1] = rco == 0 ? 1 : rco == 6 ? 3 : 2;
  // the odd result wins when its Ritz value (an upper bound of the odd
  // sector's lowest eigenvalue) lies below the even result
  const bool odd = rco != 6 && (!c.sector_state[0] || Eo < Ee);
  if (!odd) SBDX_HIP(hipMemcpyAsync(c.W, We, sizeof(double) * N, hipMemcpyDeviceToDevice, c.stream));
  SBDX_HIP(hipStreamSynchronize(c.stream));
  c.sector_chosen = odd ? 2 : 1;
  if (energy) *energy = odd ? Eo : Ee;
  if (n_steps) *n_steps = se + so;
  return odd ? rco : rce;
}

}  // namespace sbdx
