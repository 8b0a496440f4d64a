// Single-root Davidson on the device, following chemistry/tpb/davidson.h:667-930
// (and chemistry/basic/davidson_thrust.h:85-373): per step one sigma, the new
// column of the subspace matrix, a Rayleigh-Ritz eigensolve, W and the
// residual R = E0*W - H*W, both normalized; stop when |R| < eps; otherwise
// the diagonal preconditioner R/(E0-d) (1e-12 guard), modified Gram-Schmidt
// and normalization give the next basis vector; restart from W every nb.
//
// Every reduction is a fixed two-level tree (per-block partial sums in a
// fixed grid, then a fixed-order sum), so results are reproducible run to
// run.  The subspace eigensolve runs on the device (cyclic Jacobi, ascending
// order), so only |R| and E0 travel to the host once per step.
#include <chrono>
#include <cmath>

#include "common.h"

namespace sbdx {

constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;

struct VecList {
  const double *p[kNbMax];
};

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

// part[b*nv + k] = sum over this block's slice of X_k . y
__global__ void __launch_bounds__(kRedThreads) k_dots(VecList X, int nv, const double *__restrict__ y,
                                                      size_t n, double *__restrict__ part) {
  __shared__ double sh[kRedThreads];
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int k = 0; k < nv; k++) {
    const double *xk = X.p[k];
    double s = 0.0;
    for (size_t i = i0; i < n; i += stride) s += xk[i] * y[i];
    const double r = block_sum(s, sh);
    if (threadIdx.x == 0) part[(size_t)blockIdx.x * nv + k] = r;
  }
}

// Fixed-order sum of the partials of column k; written to H[k + ld*col] and
// H[col + ld*k] (the symmetric subspace matrix, column-major).
__global__ void k_finish_hcol(int nv, int nblk, const double *__restrict__ part, double *H, int ld,
                              int col) {
  const int k = threadIdx.x;
  if (k >= nv) return;
  double s = 0.0;
  for (int b = 0; b < nblk; b++) s += part[(size_t)b * nv + k];
  H[k + (size_t)ld * col] = s;
  H[col + (size_t)ld * k] = s;
}

// Cyclic Jacobi eigensolver for the leading n x n block of H (ld), one
// thread; eigenvalues ascending into E, eigenvectors (columns) into U (ld).
__global__ void k_syev(int n, int ld, const double *__restrict__ H, double *__restrict__ U,
                       double *__restrict__ E) {
  extern __shared__ double sm[];
  if (threadIdx.x != 0) return;
  double *A = sm;          // n x n
  double *V = sm + n * n;  // n x n
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      A[i + j * n] = H[i + (size_t)j * ld];
      V[i + j * n] = (i == j) ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        const double a = A[i + j * n];
        tot += a * a;
        if (i != j) off += a * a;
      }
    if (off <= 1e-32 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        const double apq = A[p + q * n];
        if (apq == 0.0) continue;
        const double app = A[p + p * n], aqq = A[q + q * n];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
        for (int k = 0; k < n; k++) {
          const double akp = A[k + p * n], akq = A[k + q * n];
          A[k + p * n] = cs * akp - sn * akq;
          A[k + q * n] = sn * akp + cs * akq;
        }
        for (int k = 0; k < n; k++) {
          const double apk = A[p + k * n], aqk = A[q + k * n];
          A[p + k * n] = cs * apk - sn * aqk;
          A[q + k * n] = sn * apk + cs * aqk;
        }
        for (int k = 0; k < n; k++) {
          const double vkp = V[k + p * n], vkq = V[k + q * n];
          V[k + p * n] = cs * vkp - sn * vkq;
          V[k + q * n] = sn * vkp + cs * vkq;
        }
      }
  }
  // selection sort of the eigenpairs, ascending
  int idx[kNbMax];
  for (int i = 0; i < n; i++) idx[i] = i;
  for (int i = 0; i < n; i++) {
    int best = i;
    for (int j = i + 1; j < n; j++)
      if (A[idx[j] + idx[j] * n] < A[idx[best] + idx[best] * n]) best = j;
    const int tmp = idx[i];
    idx[i] = idx[best];
    idx[best] = tmp;
  }
  for (int j = 0; j < n; j++) {
    E[j] = A[idx[j] + idx[j] * n];
    for (int i = 0; i < n; i++) U[i + (size_t)j * ld] = V[i + idx[j] * n];
  }
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
    double w = Cv.p[0][i] * U[0];
    double r = HCv.p[0][i] * (-U[0]);
    for (int k = 1; k < nv; k++) {
      w += Cv.p[k][i] * U[k];
      r += HCv.p[k][i] * (-U[k]);
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

__device__ __forceinline__ double sum_part(const double *part, int nblk, int stride, int off) {
  double s = 0.0;
  for (int b = 0; b < nblk; b++) s += part[(size_t)b * stride + off];
  return s;
}

// W /= |W| ; R /= |R| ; scal[0] = |R|, scal[1] = |W|.
__global__ void __launch_bounds__(kRedThreads) k_scale2(double *__restrict__ W, double *__restrict__ R,
                                                        size_t n, const double *__restrict__ part,
                                                        int nblk, double *__restrict__ scal) {
  __shared__ double f[2];
  if (threadIdx.x == 0) {
    const double nw = sqrt(sum_part(part, nblk, 2, 0));
    const double nr = sqrt(sum_part(part, nblk, 2, 1));
    f[0] = 1.0 / nw;
    f[1] = 1.0 / nr;
    if (blockIdx.x == 0) {
      scal[0] = nr;
      scal[1] = nw;
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
  if (threadIdx.x == 0) ol = -sum_part(part, nblk, 1, 0);
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
  if (threadIdx.x == 0) f = 1.0 / sqrt(sum_part(part, nblk, 1, 0));
  __syncthreads();
  const double g = f;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) c[i] *= g;
}

__global__ void k_finish_scalar(int nblk, const double *__restrict__ part, double *out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *out = sum_part(part, nblk, 1, 0);
}

static void dots(Context &c, const double *const *X, int nv, const double *y, double *part) {
  VecList L{};
  for (int k = 0; k < nv; k++) L.p[k] = X[k];
  hipLaunchKernelGGL(k_dots, dim3(kRedBlocks), dim3(kRedThreads), 0, c.stream, L, nv, y, c.N, part);
  check_launch("k_dots");
}

double device_dot(Context &c, const double *x, const double *y, size_t n) {
  (void)n;
  const double *X[1] = {x};
  dots(c, X, 1, y, c.part);
  double *scal = c.dsmall + kScalarOffset + 8;
  hipLaunchKernelGGL(k_finish_scalar, dim3(1), dim3(1), 0, c.stream, kRedBlocks, c.part, scal);
  check_launch("k_finish_scalar");
  double r = 0.0;
  SBDX_HIP(hipMemcpyAsync(&r, scal, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  SBDX_HIP(hipStreamSynchronize(c.stream));
  return r;
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
      sigma_alpha_first(c, c.C[ib], c.HC[ib]);
      {
        ProfScope ps(c, "dav_subspace");
        dots(c, c.C.data(), ib + 1, c.HC[ib], c.part);
        hipLaunchKernelGGL(k_finish_hcol, dim3(1), dim3(kNbMax), 0, c.stream, ib + 1, kRedBlocks,
                           c.part, H, ld, ib);
        check_launch("k_finish_hcol");
        const size_t shm = sizeof(double) * 2 * (ib + 1) * (ib + 1);
        hipLaunchKernelGGL(k_syev, dim3(1), dim3(64), shm, c.stream, ib + 1, ld, H, U, E);
        check_launch("k_syev");
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
      if (nr < eps) {
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

}  // namespace sbdx
