// C ABI of libsbdx (declared in include/sbdx.h): handle lifecycle, the
// boundary conversion between the reference's interleaved determinant basis
// and the device's alpha-first basis, profiling, and error reporting.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "../../include/sbdx.h"
#include "common.h"

namespace sbdx {

static thread_local std::string g_err;

void *dalloc(Context &c, size_t bytes) {
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, bytes ? bytes : 8);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error(SBDX_ERR_ALLOC, "hipMalloc(" + std::to_string(bytes) + " bytes) failed: " +
                                    hipGetErrorString(e));
  }
  c.allocs.push_back(p);
  return p;
}

void dfree(Context &c, void *p) {
  if (!p) return;
  for (size_t i = 0; i < c.allocs.size(); i++)
    if (c.allocs[i] == p) {
      (void)hipFree(p);
      c.allocs.erase(c.allocs.begin() + (long)i);
      return;
    }
}

void dfree_all(Context &c) {
  for (void *p : c.allocs) (void)hipFree(p);
  c.allocs.clear();
}

TmpBuf::TmpBuf(size_t bytes) {
  hipError_t e = hipMalloc(&p, bytes ? bytes : 8);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    p = nullptr;
    throw Error(SBDX_ERR_ALLOC, "hipMalloc(" + std::to_string(bytes) + " bytes) failed");
  }
}
TmpBuf::~TmpBuf() {
  if (p) (void)hipFree(p);
}

void check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw Error(SBDX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

ProfScope::ProfScope(Context &cx, const char *nm) : c(cx), name(nm) {
  if (!c.prof || c.capturing) return;
  SBDX_HIP(hipEventCreate(&b));
  SBDX_HIP(hipEventCreate(&e));
  SBDX_HIP(hipEventRecord(b, c.stream));
}
ProfScope::~ProfScope() {
  if (!b) return;
  if (hipEventRecord(e, c.stream) != hipSuccess) return;
  auto it = c.prof_slots.find(name);
  if (it == c.prof_slots.end()) {
    c.prof_order.push_back(name);
    it = c.prof_slots.emplace(name, ProfSlot{}).first;
  }
  it->second.ev.emplace_back(b, e);
  it->second.launches++;
}

static void prof_collect(Context &c) {
  SBDX_HIP(hipStreamSynchronize(c.stream));
  for (auto &kv : c.prof_slots) {
    for (auto &pr : kv.second.ev) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) kv.second.total_ms += ms;
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
    kv.second.ev.clear();
  }
}

static void prof_clear(Context &c) {
  for (auto &kv : c.prof_slots)
    for (auto &pr : kv.second.ev) {
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
  c.prof_slots.clear();
  c.prof_order.clear();
}

static size_t i2_size(int L) {
  const size_t B = (size_t)L * (L + 1) / 2;
  return B * (B + 1) / 2;
}
static inline size_t tri(size_t a, size_t b) {
  const size_t mx = a > b ? a : b, mn = a > b ? b : a;
  return mx * (mx + 1) / 2 + mn;
}

static std::vector<uint64_t> unpack_words(const uint64_t *w, size_t n, size_t bl, int L) {
  const size_t nw = (L + bl - 1) / bl;
  std::vector<uint64_t> out(n);
  const uint64_t lowmask = bl >= 64 ? ~0ull : ((1ull << bl) - 1ull);
  const uint64_t lmask = L >= 64 ? ~0ull : ((1ull << L) - 1ull);
  for (size_t i = 0; i < n; i++) {
    uint64_t m = 0;
    for (size_t k = 0; k < nw; k++) m |= (w[i * nw + k] & lowmask) << (k * bl);
    out[i] = m & lmask;
  }
  return out;
}

static void upload(Context &c, double *d, const double *h) {
  SBDX_HIP(hipMemcpyAsync(d, h, sizeof(double) * c.N, hipMemcpyHostToDevice, c.stream));
}
static void download(Context &c, double *h, const double *d) {
  SBDX_HIP(hipMemcpyAsync(h, d, sizeof(double) * c.N, hipMemcpyDeviceToHost, c.stream));
  SBDX_HIP(hipStreamSynchronize(c.stream));
}

static Context *create_core(int L, std::vector<uint64_t> as, std::vector<uint64_t> bs, double I0,
                            std::vector<double> h1, std::vector<double> eri);

static Context *create(int L, size_t bl, const uint64_t *aw, size_t na, const uint64_t *bw, size_t nb,
                       double I0, const double *I1, const double *I2, size_t nI2) {
  if (L < 1 || L > 64) throw Error(SBDX_ERR_UNSUPPORTED, "norb must be in [1, 64]");
  if (bl < 1 || bl > 64) throw Error(SBDX_ERR_ARG, "bit_length must be in [1, 64]");
  if (!aw || !bw || !I1 || !I2) throw Error(SBDX_ERR_ARG, "null input pointer");
  if (na < 1 || nb < 1) throw Error(SBDX_ERR_ARG, "empty string set");
  if (na > 0xffffffffull || nb > 0xffffffffull)
    throw Error(SBDX_ERR_UNSUPPORTED, "string sets are limited to 2^32-1 strings");
  if (nI2 != i2_size(L))
    throw Error(SBDX_ERR_ARG, "n_I2 must be B(B+1)/2 with B = L(L+1)/2 (" + std::to_string(i2_size(L)) + ")");
  // integrals: spatial one-body from the alpha block of I1, dense (pq|rs)
  const size_t K = (size_t)L * L;
  std::vector<double> h1(K), eri(K * K);
  for (int p = 0; p < L; p++)
    for (int q = 0; q < L; q++) h1[p * L + q] = I1[(size_t)(2 * p) * (2 * L) + 2 * q];
  for (int p = 0; p < L; p++)
    for (int q = 0; q < L; q++) {
      const size_t pq = tri(p, q);
      for (int r = 0; r < L; r++)
        for (int s = 0; s < L; s++) eri[(p * L + q) * K + r * L + s] = I2[tri(pq, tri(r, s))];
    }
  return create_core(L, unpack_words(aw, na, bl, L), unpack_words(bw, nb, bl, L), I0, std::move(h1),
                     std::move(eri));
}

// A context from unpacked strings and dense spatial integrals (kept on the
// host too, for sub-contexts of the coarse start).
static Context *create_core(int L, std::vector<uint64_t> as, std::vector<uint64_t> bs, double I0,
                            std::vector<double> h1, std::vector<double> eri) {
  const size_t na = as.size(), nb = bs.size();
  auto ctx = std::make_unique<Context>();
  Context &c = *ctx;
  try {
    SBDX_HIP(hipGetDevice(&c.device));
    int lds = 0, ncu = 0;
    SBDX_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c.device));
    SBDX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c.device));
    c.lds_limit = (size_t)lds;
    // SBDX_LDS_LIMIT=<bytes> lowers the per-block LDS budget the kernels plan
    // for, which selects the row-parts and CSR fallbacks on small instances.
    if (const char *ev = std::getenv("SBDX_LDS_LIMIT")) {
      const long v = std::atol(ev);
      if (v > 0 && (size_t)v < c.lds_limit) c.lds_limit = (size_t)v;
    }
    c.n_cu = ncu;
    SBDX_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    c.L = L;
    c.K = L * L;
    c.na = na;
    c.nb = nb;
    c.N = na * nb;
    c.I0 = I0;
    const size_t K = (size_t)c.K;
    c.h1 = (double *)dalloc(c, sizeof(double) * K);
    c.eri = (double *)dalloc(c, sizeof(double) * K * K);
    SBDX_HIP(hipMemcpyAsync(c.h1, h1.data(), sizeof(double) * K, hipMemcpyHostToDevice, c.stream));
    SBDX_HIP(hipMemcpyAsync(c.eri, eri.data(), sizeof(double) * K * K, hipMemcpyHostToDevice, c.stream));
    c.eri_rs_sym = true;
    for (size_t pq = 0; pq < K && c.eri_rs_sym; pq++)
      for (int a = 1; a < L && c.eri_rs_sym; a++)
        for (int b = 0; b < a; b++)
          if (eri[pq * K + (size_t)a * L + b] != eri[pq * K + (size_t)b * L + a]) {
            c.eri_rs_sym = false;
            break;
          }
    c.Kt = L * (L + 1) / 2;
    {
      std::vector<uint16_t> tri((size_t)c.Kt);
      for (int a = 0; a < L; a++)
        for (int b = 0; b <= a; b++) tri[(size_t)a * (a + 1) / 2 + b] = (uint16_t)(a * L + b);
      c.tri_rs = (uint16_t *)dalloc(c, sizeof(uint16_t) * tri.size());
      SBDX_HIP(hipMemcpyAsync(c.tri_rs, tri.data(), sizeof(uint16_t) * tri.size(), hipMemcpyHostToDevice,
                              c.stream));
      SBDX_HIP(hipStreamSynchronize(c.stream));
    }
    build_spin_tables(c, c.A, as);
    build_singles_ell(c, c.A);
    if (as == bs) {
      c.B = c.A;
      c.B.alias = true;
    } else {
      build_spin_tables(c, c.B, bs);
      build_singles_ell(c, c.B);
    }
    c.hii = (double *)dalloc(c, sizeof(double) * c.N);
    compute_diag(c);
    c.vin = (double *)dalloc(c, sizeof(double) * c.N);
    c.vout = (double *)dalloc(c, sizeof(double) * c.N);
    c.T1 = (double *)dalloc(c, sizeof(double) * c.N);
    c.T2 = (double *)dalloc(c, sizeof(double) * c.N);
    c.W = (double *)dalloc(c, sizeof(double) * c.N);
    c.part_cap = 1024 * (size_t)kNbMax;
    c.part = (double *)dalloc(c, sizeof(double) * c.part_cap);
    c.dsmall = (double *)dalloc(c, sizeof(double) * kSmallDoubles);
    SBDX_HIP(hipMemsetAsync(c.dsmall, 0, sizeof(double) * kSmallDoubles, c.stream));
    SBDX_HIP(hipStreamSynchronize(c.stream));
    c.h_astr = std::move(as);
    c.h_bstr = std::move(bs);
    c.h_h1 = std::move(h1);
    c.h_eri = std::move(eri);
  } catch (...) {
    dfree_all(c);
    if (c.stream) (void)hipStreamDestroy(c.stream);
    throw;
  }
  return ctx.release();
}

// The problem restricted to the first m alpha and m beta strings (same
// integrals), for the coarse start of the thick-restart solves.
Context *make_sub_context(const Context &c, size_t m) {
  if (m < 1 || m > c.h_astr.size() || m > c.h_bstr.size()) throw Error(1, "sub-context size out of range");
  std::vector<uint64_t> as(c.h_astr.begin(), c.h_astr.begin() + m);
  std::vector<uint64_t> bs(c.h_bstr.begin(), c.h_bstr.begin() + m);
  Context *s = create_core(c.L, std::move(as), std::move(bs), c.I0, c.h_h1, c.h_eri);
  s->lds_limit = c.lds_limit;
  s->mixed_form = c.mixed_form;
  return s;
}

static void destroy(Context *c) {
  if (!c) return;
  if (c->coarse_sub) destroy(c->coarse_sub);
  (void)hipStreamSynchronize(c->stream);
  prof_clear(*c);
  graphs_release(*c);
  same_dense_release(*c);
  if (c->hpin) (void)hipHostFree(c->hpin);
  dfree_all(*c);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

void destroy_context(Context *c) { destroy(c); }

}  // namespace sbdx

using namespace sbdx;

#define SBDX_GUARD_BEGIN try {
#define SBDX_GUARD_END                                        \
  }                                                           \
  catch (const sbdx::Error &e) {                              \
    g_err = e.what();                                         \
    return e.code;                                            \
  }                                                           \
  catch (const std::bad_alloc &) {                            \
    g_err = "host allocation failed";                         \
    return SBDX_ERR_ALLOC;                                    \
  }                                                           \
  catch (const std::exception &e) {                           \
    g_err = e.what();                                         \
    return SBDX_ERR_ARG;                                      \
  }                                                           \
  catch (...) {                                               \
    g_err = "unknown error";                                  \
    return SBDX_ERR_ARG;                                      \
  }

static Context &ctx_of(sbdx_handle_t h) {
  if (!h) throw Error(SBDX_ERR_ARG, "null handle");
  return *reinterpret_cast<Context *>(h);
}

extern "C" {

const char *sbdx_last_error(void) { return g_err.c_str(); }
int sbdx_version(void) { return 100; }

int sbdx_create(sbdx_handle_t *out, int norb, size_t bit_length, const uint64_t *adet_words,
                size_t n_adet, const uint64_t *bdet_words, size_t n_bdet, double I0,
                const double *I1, const double *I2, size_t n_I2) {
  SBDX_GUARD_BEGIN
  if (!out) throw Error(SBDX_ERR_ARG, "null output handle pointer");
  *out = nullptr;
  Context *c = create(norb, bit_length, adet_words, n_adet, bdet_words, n_bdet, I0, I1, I2, n_I2);
  *out = reinterpret_cast<sbdx_handle_t>(c);
  g_err.clear();
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_destroy(sbdx_handle_t h) {
  SBDX_GUARD_BEGIN
  destroy(reinterpret_cast<Context *>(h));
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_dims(sbdx_handle_t h, size_t *na, size_t *nb, size_t *nd) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (na) *na = c.na;
  if (nb) *nb = c.nb;
  if (nd) *nd = c.N;
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_diag(sbdx_handle_t h, double *hii) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (!hii) throw Error(SBDX_ERR_ARG, "null hii");
  download(c, hii, c.hii);
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_sigma_device(sbdx_handle_t h, const double *d_in, double *d_out, int accumulate) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (!d_in || !d_out) throw Error(SBDX_ERR_ARG, "null vector");
  sign_convert(c, d_in, c.vin);
  sigma_alpha_first(c, c.vin, c.vout);
  if (accumulate)
    sign_convert_add(c, c.vout, d_out);
  else
    sign_convert(c, c.vout, d_out);
  SBDX_HIP(hipStreamSynchronize(c.stream));
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_sigma(sbdx_handle_t h, const double *Wk, double *Wb, int accumulate) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (!Wk || !Wb) throw Error(SBDX_ERR_ARG, "null vector");
  upload(c, c.vout, Wk);
  sign_convert(c, c.vout, c.vin);
  sigma_alpha_first(c, c.vin, c.vout);
  if (accumulate) {
    upload(c, c.vin, Wb);
    sign_convert_add(c, c.vout, c.vin);
    download(c, Wb, c.vin);
  } else {
    sign_convert(c, c.vout, c.vin);
    download(c, Wb, c.vin);
  }
  return SBDX_OK;
  SBDX_GUARD_END
}

static int sigma_block_host(sbdx_handle_t h, int k, const double *Wk, double *Wb, int accumulate,
                            bool fp32) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (!Wk || !Wb) throw Error(SBDX_ERR_ARG, "null vector");
  if (k < 1 || k > 64) throw Error(SBDX_ERR_ARG, "k must be in [1, 64]");
  const size_t N = c.N;
  TmpBuf xb(sizeof(double) * N * k), yb(sizeof(double) * N * k);
  double *x = (double *)xb.p, *y = (double *)yb.p;
  for (int v = 0; v < k; v++) {
    upload(c, c.vout, Wk + (size_t)v * N);
    sign_convert(c, c.vout, x + (size_t)v * N);
  }
  if (fp32)
    sigma_block_f32(c, k, x, y);
  else
    sigma_block(c, k, x, y);
  for (int v = 0; v < k; v++) {
    if (accumulate) {
      upload(c, c.vin, Wb + (size_t)v * N);
      sign_convert_add(c, y + (size_t)v * N, c.vin);
    } else {
      sign_convert(c, y + (size_t)v * N, c.vin);
    }
    download(c, Wb + (size_t)v * N, c.vin);
  }
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_sigma_block(sbdx_handle_t h, int k, const double *Wk, double *Wb, int accumulate) {
  return sigma_block_host(h, k, Wk, Wb, accumulate, false);
}

int sbdx_sigma_block_fp32(sbdx_handle_t h, int k, const double *Wk, double *Wb) {
  return sigma_block_host(h, k, Wk, Wb, 0, true);
}

int sbdx_davidson_block(sbdx_handle_t h, int nroots, int block, double *W, int max_it, double eps,
                        double max_time, double *energies, int *n_products) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (!W || !energies) throw Error(SBDX_ERR_ARG, "null argument");
  if (nroots < 1 || block < nroots || block > 16)
    throw Error(SBDX_ERR_ARG, "need 1 <= nroots <= block <= 16");
  const size_t N = c.N;
  TmpBuf w0(sizeof(double) * N * nroots), w1(sizeof(double) * N * nroots);
  double *a = (double *)w0.p, *b = (double *)w1.p;
  for (int j = 0; j < nroots; j++) {
    upload(c, c.vin, W + (size_t)j * N);
    sign_convert(c, c.vin, a + (size_t)j * N);
  }
  const int rc = davidson_block(c, nroots, block, max_it, eps, max_time, c.precision == 1, a, b, energies,
                                n_products);
  for (int j = 0; j < nroots; j++) {
    sign_convert(c, b + (size_t)j * N, c.vin);
    download(c, W + (size_t)j * N, c.vin);
  }
  if (rc != 0) throw Error(SBDX_ERR_NOT_CONVERGED, "davidson_block: not converged");
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_set_block_basis(sbdx_handle_t h, int m) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (m != 0 && (m < 3 || m > kNbMax)) throw Error(SBDX_ERR_ARG, "block basis size must be 0 or in [3, 64]");
  c.blk_m_req = m;
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_block_stats(sbdx_handle_t h, int *basis, int *keep, int *restarts, double *resid, int *parity,
                     int nroots) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (basis) *basis = c.blk_m_used;
  if (keep) *keep = c.blk_keep_used;
  if (restarts) *restarts = c.blk_restarts;
  if (resid)
    for (int j = 0; j < nroots; j++) resid[j] = j < (int)c.blk_resid.size() ? c.blk_resid[j] : 0.0;
  if (parity)
    for (int j = 0; j < nroots; j++) parity[j] = j < (int)c.blk_par.size() ? c.blk_par[j] : 0;
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_set_vector(sbdx_handle_t h, const double *W) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (!W) throw Error(SBDX_ERR_ARG, "null vector");
  upload(c, c.vin, W);
  sign_convert(c, c.vin, c.W);
  SBDX_HIP(hipStreamSynchronize(c.stream));
  c.have_vector = true;
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_get_vector(sbdx_handle_t h, double *W) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (!W) throw Error(SBDX_ERR_ARG, "null vector");
  if (!c.have_vector) throw Error(SBDX_ERR_ARG, "no vector set");
  sign_convert(c, c.W, c.vout);
  download(c, W, c.vout);
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_solve(sbdx_handle_t h, int max_it, int nb, double eps, double max_time, double *energy,
               int *n_steps) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (!c.have_vector) throw Error(SBDX_ERR_ARG, "no start vector (sbdx_set_vector)");
  if (max_it < 1) throw Error(SBDX_ERR_ARG, "max_it must be >= 1");
  const int rc = c.restart_keep == 0
                     ? davidson_solve(c, max_it, nb, eps, max_time, energy, n_steps)
                     : davidson_solve_sectors(c, max_it * nb, nb, c.restart_keep, eps, max_time,
                                              c.precision == SBDX_PREC_MIXED, energy, n_steps);
  if (rc != 0) g_err = "Davidson did not reach eps";
  return rc;
  SBDX_GUARD_END
}

int sbdx_davidson(sbdx_handle_t h, double *W, int max_it, int nb, double eps, double max_time,
                  double *energy, int *n_steps) {
  int rc = sbdx_set_vector(h, W);
  if (rc != SBDX_OK) return rc;
  rc = sbdx_solve(h, max_it, nb, eps, max_time, energy, n_steps);
  if (rc != SBDX_OK && rc != SBDX_ERR_NOT_CONVERGED) return rc;
  const int rc2 = sbdx_get_vector(h, W);
  return rc2 != SBDX_OK ? rc2 : rc;
}

int sbdx_set_precision(sbdx_handle_t h, int mode) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (mode != SBDX_PREC_FP64 && mode != SBDX_PREC_MIXED) throw Error(SBDX_ERR_ARG, "unknown precision mode");
  if (mode == SBDX_PREC_MIXED) ensure_f32(c);
  c.precision = mode;
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_set_restart(sbdx_handle_t h, int keep) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (keep < 0 || keep > 8) throw Error(SBDX_ERR_ARG, "keep must be in [0, 8]");
  c.restart_keep = keep;
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_same_spin_form(sbdx_handle_t h, int precision, int *form, float *ms_sparse,
                        float *ms_dense) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (precision != 0 && precision != 1) throw Error(SBDX_ERR_ARG, "precision must be 0 or 1");
  if (form) *form = c.same_mode[precision];
  if (ms_sparse) *ms_sparse = c.same_ms[precision][0];
  if (ms_dense) *ms_dense = c.same_ms[precision][1];
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_fp64_switch_step(sbdx_handle_t h, int *step) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (!step) throw Error(SBDX_ERR_ARG, "null pointer");
  *step = c.fp64_switch_step;
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_split_gemm(sbdx_handle_t h, int *mode, float *ms_sgemm, float *ms_split) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (mode) *mode = c.split_mode;
  if (ms_sgemm) *ms_sgemm = c.split_ms[0];
  if (ms_split) *ms_split = c.split_ms[1];
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_set_mixed_form(sbdx_handle_t h, int form) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (form < 0 || form > 2) throw Error(SBDX_ERR_ARG, "mixed form must be 0, 1 or 2");
  c.mixed_form = form;
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_mixed_staging(sbdx_handle_t h, int precision, int *width, int *slots_per_thread,
                       float *ms, int *multi) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (precision < 0 || precision > 3) throw Error(SBDX_ERR_ARG, "precision must be 0, 1, 2 or 3");
  // the row-parts kernel (no staging width fits) reports width -1 and its
  // own tuned slots per thread
  const bool parts = c.mixed_eb[precision] <= 0 && c.parts_nt[precision] > 0;
  if (width) *width = parts ? -1 : c.mixed_eb[precision];
  if (slots_per_thread) *slots_per_thread = parts ? c.parts_nt[precision] : c.mixed_nt[precision];
  if (ms) *ms = c.mixed_best_ms[precision];
  if (multi) *multi = precision < 2 ? c.multi_mode[precision] : -1;
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_symmetric_eig(int n, const double *A, double *V, double *E) {
  SBDX_GUARD_BEGIN
  if (n < 1 || n > 4096) throw Error(SBDX_ERR_ARG, "n must be in [1, 4096]");
  if (!A || !V || !E) throw Error(SBDX_ERR_ARG, "null argument");
  std::vector<double> a(A, A + (size_t)n * n);
  host_syev(n, a.data(), V, E);
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_symmetric_geig(int n, const double *H, const double *S, double *Y, double *E) {
  SBDX_GUARD_BEGIN
  if (n < 1 || n > 4096) throw Error(SBDX_ERR_ARG, "n must be in [1, 4096]");
  if (!H || !S || !Y || !E) throw Error(SBDX_ERR_ARG, "null argument");
  host_sygv(n, H, S, n, Y, E);
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_set_symmetry(sbdx_handle_t h, int mode) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (mode < SBDX_SYM_OFF || mode > SBDX_SYM_LOWEST)
    throw Error(SBDX_ERR_ARG, "spin-flip symmetry mode must be 0 (off), 1 (even), 2 (odd) or 3 (lowest of both)");
  if (mode && !sigma_sym_available(c))
    throw Error(SBDX_ERR_ARG, "spin-flip symmetry needs equal alpha and beta string sets (and memory for the "
                              "dense same-spin block)");
  c.symmetric = mode;
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_sector_stats(sbdx_handle_t h, double *energies, int *products, int *states, int *chosen) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  for (int k = 0; k < 2; k++) {
    if (energies) energies[k] = c.sector_E[k];
    if (products) products[k] = c.sector_products[k];
    if (states) states[k] = c.sector_state[k];
  }
  if (chosen) *chosen = c.sector_chosen;
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_sigma_sym(sbdx_handle_t h, const double *Wk, double *Wb, int fp32) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (!Wk || !Wb) throw Error(SBDX_ERR_ARG, "null vector");
  if (!sigma_sym_available(c)) throw Error(SBDX_ERR_ARG, "spin-flip symmetry not available for this handle");
  upload(c, c.vout, Wk);
  sign_convert(c, c.vout, c.vin);
  // the sector is the vector's own: even (x = x^T) or odd (x = -x^T)
  int par = 1;
  if (sym_defect(c, c.vin, 1) > 1e-6) {
    if (sym_defect(c, c.vin, -1) > 1e-6)
      throw Error(SBDX_ERR_ARG, "vector lies in neither spin-flip sector (x = x^T or x = -x^T)");
    par = -1;
  }
  const int saved = c.sym_par;
  c.sym_par = par;
  try {
    if (fp32)
      sigma_sym_f32(c, c.vin, c.vout);
    else
      sigma_sym(c, c.vin, c.vout);
  } catch (...) {
    c.sym_par = saved;
    throw;
  }
  c.sym_par = saved;
  sign_convert(c, c.vout, c.vin);
  download(c, Wb, c.vin);
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_mlp_apply(sbdx_handle_t h, const double *R, double E, double *out) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (!R || !out) throw Error(SBDX_ERR_ARG, "null vector");
  if (!c.mlp.valid) throw Error(SBDX_ERR_ARG, "no preconditioner weights set");
  upload(c, c.vin, R);
  double *dE = c.dsmall + kScalarOffset + 16;
  SBDX_HIP(hipMemcpyAsync(dE, &E, sizeof(double), hipMemcpyHostToDevice, c.stream));
  mlp_precond(c, c.vin, dE, c.vout);
  download(c, out, c.vout);
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_set_coarse_start(sbdx_handle_t h, size_t m, double eps) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (m > 0 && (m >= c.na || m >= c.nb)) throw Error(SBDX_ERR_ARG, "coarse start needs m below both string counts");
  if (!(eps > 0.0)) throw Error(SBDX_ERR_ARG, "eps must be positive");
  c.coarse_m = m;
  c.coarse_eps = eps;
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_coarse_stats(sbdx_handle_t h, int *products, double *energy, double *seconds) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (products) *products = c.coarse_products;
  if (energy) *energy = c.coarse_energy;
  if (seconds) *seconds = c.coarse_seconds;
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_set_graphs(sbdx_handle_t h, int on) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (!on) graphs_release(c);
  c.use_graphs = on ? 1 : 0;
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_set_verbose(sbdx_handle_t h, int level) {
  SBDX_GUARD_BEGIN
  ctx_of(h).verbose = level;
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_sigma_fp32(sbdx_handle_t h, const double *Wk, double *Wb) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (!Wk || !Wb) throw Error(SBDX_ERR_ARG, "null vector");
  upload(c, c.vout, Wk);
  sign_convert(c, c.vout, c.vin);
  sigma_alpha_first_f32(c, c.vin, c.vout);
  sign_convert(c, c.vout, c.vin);
  download(c, Wb, c.vin);
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_mlp_nparam(void) { return mlp_nparam(); }
int sbdx_mlp_nfeat(void) { return mlp_nfeat(); }

int sbdx_mlp_set_weights(sbdx_handle_t h, const float *params, size_t n) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (!params || n != (size_t)mlp_nparam()) throw Error(SBDX_ERR_ARG, "mlp: wrong parameter count");
  mlp_set(c, params);
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_mlp_get_weights(sbdx_handle_t h, float *params, size_t n) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (!params || n != (size_t)mlp_nparam()) throw Error(SBDX_ERR_ARG, "mlp: wrong parameter count");
  mlp_get(c, params);
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_mlp_forward(sbdx_handle_t h, const uint16_t *feat, size_t n, float *out) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (!feat || !out || n == 0) throw Error(SBDX_ERR_ARG, "mlp: empty input");
  mlp_ensure(c);
  const size_t nf = (size_t)mlp_nfeat();
  TmpBuf df(sizeof(uint16_t) * n * nf), dy(sizeof(float) * n);
  SBDX_HIP(hipMemcpyAsync(df.p, feat, sizeof(uint16_t) * n * nf, hipMemcpyHostToDevice, c.stream));
  mlp_forward(c, (const unsigned short *)df.p, n, (float *)dy.p);
  SBDX_HIP(hipMemcpyAsync(out, dy.p, sizeof(float) * n, hipMemcpyDeviceToHost, c.stream));
  SBDX_HIP(hipStreamSynchronize(c.stream));
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_mlp_gradient(sbdx_handle_t h, const uint16_t *feat, const float *target, size_t n, float *grad,
                      double *sse) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (!feat || !target || !grad || n == 0) throw Error(SBDX_ERR_ARG, "mlp: empty input");
  mlp_ensure(c);
  const size_t nf = (size_t)mlp_nfeat();
  TmpBuf df(sizeof(uint16_t) * n * nf), dt(sizeof(float) * n), di(sizeof(uint32_t) * n);
  std::vector<uint32_t> idx(n);
  for (size_t i = 0; i < n; i++) idx[i] = (uint32_t)i;
  SBDX_HIP(hipMemcpyAsync(df.p, feat, sizeof(uint16_t) * n * nf, hipMemcpyHostToDevice, c.stream));
  SBDX_HIP(hipMemcpyAsync(dt.p, target, sizeof(float) * n, hipMemcpyHostToDevice, c.stream));
  SBDX_HIP(hipMemcpyAsync(di.p, idx.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, c.stream));
  const double s = mlp_gradient(c, (const unsigned short *)df.p, (const float *)dt.p,
                                (const uint32_t *)di.p, n);
  if (sse) *sse = s;
  SBDX_HIP(hipMemcpyAsync(grad, c.mlp.grad, sizeof(float) * mlp_nparam(), hipMemcpyDeviceToHost, c.stream));
  SBDX_HIP(hipStreamSynchronize(c.stream));
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_mlp_train(sbdx_handle_t h, const uint16_t *feat, const float *target, size_t n, int epochs,
                   int batch, float lr, uint64_t seed, double *loss) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (!feat || !target || n == 0 || epochs < 1 || batch < 1) throw Error(SBDX_ERR_ARG, "mlp: bad training arguments");
  mlp_ensure(c);
  const size_t nf = (size_t)mlp_nfeat();
  TmpBuf df(sizeof(uint16_t) * n * nf), dt(sizeof(float) * n);
  SBDX_HIP(hipMemcpyAsync(df.p, feat, sizeof(uint16_t) * n * nf, hipMemcpyHostToDevice, c.stream));
  SBDX_HIP(hipMemcpyAsync(dt.p, target, sizeof(float) * n, hipMemcpyHostToDevice, c.stream));
  mlp_train(c, (const unsigned short *)df.p, (const float *)dt.p, n, epochs, batch, lr, seed, loss);
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_mlp_record(sbdx_handle_t h, int cap) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (cap < 0 || cap > 4096) throw Error(SBDX_ERR_ARG, "mlp: record cap must be in [0, 4096]");
  c.mlp.record_cap = cap;
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_mlp_collect(sbdx_handle_t h, size_t max_pairs, uint64_t seed, uint16_t *feat, float *target,
                     size_t *n_pairs) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (!feat || !target || !n_pairs) throw Error(SBDX_ERR_ARG, "mlp: null output");
  *n_pairs = mlp_collect(c, max_pairs, seed, (unsigned short *)feat, target);
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_set_preconditioner(sbdx_handle_t h, int mode) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (mode != 0 && mode != 1) throw Error(SBDX_ERR_ARG, "preconditioner mode must be 0 or 1");
  if (mode == 1 && !c.mlp.valid) throw Error(SBDX_ERR_ARG, "mlp preconditioner has no weights");
  c.precond = mode;
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_precond_stats(sbdx_handle_t h, long *mlp_steps, long *diag_steps) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (mlp_steps) *mlp_steps = c.mlp.used_mlp;
  if (diag_steps) *diag_steps = c.mlp.used_diag;
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_trace(sbdx_handle_t h, double *trace, int cap, int *n) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  const int steps = (int)(c.trace.size() / 2);
  if (n) *n = steps;
  if (trace)
    for (int k = 0; k < std::min(cap, steps); k++) {
      trace[2 * k] = c.trace[2 * k];
      trace[2 * k + 1] = c.trace[2 * k + 1];
    }
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_energy(sbdx_handle_t h, const double *W, double *energy) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (!W || !energy) throw Error(SBDX_ERR_ARG, "null pointer");
  upload(c, c.vin, W);
  sign_convert(c, c.vin, c.vout);
  sigma_alpha_first(c, c.vout, c.vin);
  *energy = device_dot(c, c.vout, c.vin, c.N);
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_rdm(sbdx_handle_t h, const double *W, double *one, double *two) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (!W || !one || !two) throw Error(SBDX_ERR_ARG, "null pointer");
  const size_t L2 = (size_t)c.L * c.L;
  upload(c, c.vin, W);
  sign_convert(c, c.vin, c.vout);
  TmpBuf d1(sizeof(double) * 2 * L2), d2(sizeof(double) * 4 * L2 * L2);
  compute_rdm(c, c.vout, (double *)d1.p, (double *)d2.p);
  SBDX_HIP(hipMemcpyAsync(one, d1.p, sizeof(double) * 2 * L2, hipMemcpyDeviceToHost, c.stream));
  SBDX_HIP(hipMemcpyAsync(two, d2.p, sizeof(double) * 4 * L2 * L2, hipMemcpyDeviceToHost, c.stream));
  SBDX_HIP(hipStreamSynchronize(c.stream));
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_rdm_device_noalloc(sbdx_handle_t h, const double *d_W, double *d_one, double *d_two) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (!d_W || !d_one || !d_two) throw Error(SBDX_ERR_ARG, "null pointer");
  // c.vout and c.T1 are the handle's own buffers, allocated by sbdx_create
  sign_convert_noprof(c, d_W, c.vout);
  compute_rdm_noalloc(c, c.vout, d_one, d_two);
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_table_count(sbdx_handle_t h, int spin, int kind, size_t *count) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (!count || (spin != 0 && spin != 1) || (kind != 1 && kind != 2))
    throw Error(SBDX_ERR_ARG, "bad table selector");
  const SpinTables &t = spin ? c.B : c.A;
  *count = kind == 1 ? t.ns : t.nd;
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_table_export(sbdx_handle_t h, int spin, int kind, uint64_t *offsets, uint32_t *kets,
                      int32_t *cran) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (!offsets || !kets || !cran || (spin != 0 && spin != 1) || (kind != 1 && kind != 2))
    throw Error(SBDX_ERR_ARG, "bad table export arguments");
  const SpinTables &t = spin ? c.B : c.A;
  const size_t n = t.n;
  if (kind == 1) {
    std::vector<uint64_t> ent(t.ns);
    SBDX_HIP(hipMemcpyAsync(offsets, t.s_off, sizeof(uint64_t) * (n + 1), hipMemcpyDeviceToHost, c.stream));
    SBDX_HIP(hipMemcpyAsync(ent.data(), t.s_ent, sizeof(uint64_t) * t.ns, hipMemcpyDeviceToHost, c.stream));
    SBDX_HIP(hipStreamSynchronize(c.stream));
    for (size_t e = 0; e < t.ns; e++) {
      const uint32_t pq = (uint32_t)((ent[e] >> 32) & 0x7fffffffull);
      kets[e] = (uint32_t)(ent[e] & 0xffffffffull);
      cran[2 * e] = 2 * (int)(pq / c.L) + spin;
      cran[2 * e + 1] = 2 * (int)(pq % c.L) + spin;
    }
  } else {
    std::vector<uint32_t> orb(t.nd);
    SBDX_HIP(hipMemcpyAsync(offsets, t.d_off, sizeof(uint64_t) * (n + 1), hipMemcpyDeviceToHost, c.stream));
    SBDX_HIP(hipMemcpyAsync(kets, t.d_ket, sizeof(uint32_t) * t.nd, hipMemcpyDeviceToHost, c.stream));
    SBDX_HIP(hipMemcpyAsync(orb.data(), t.d_orb, sizeof(uint32_t) * t.nd, hipMemcpyDeviceToHost, c.stream));
    SBDX_HIP(hipStreamSynchronize(c.stream));
    for (size_t e = 0; e < t.nd; e++)
      for (int k = 0; k < 4; k++) cran[4 * e + k] = 2 * (int)((orb[e] >> (8 * k)) & 255u) + spin;
  }
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_profile(sbdx_handle_t h, int enable) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  SBDX_HIP(hipStreamSynchronize(c.stream));
  prof_clear(c);
  c.prof = enable != 0;
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_profile_count(sbdx_handle_t h, int *n) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  prof_collect(c);
  if (n) *n = (int)c.prof_order.size();
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_profile_get(sbdx_handle_t h, int idx, char *name, size_t cap, long *launches,
                     double *total_ms) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  prof_collect(c);
  if (idx < 0 || idx >= (int)c.prof_order.size()) throw Error(SBDX_ERR_ARG, "profile index out of range");
  const std::string &nm = c.prof_order[idx];
  const ProfSlot &s = c.prof_slots[nm];
  if (name && cap) {
    std::strncpy(name, nm.c_str(), cap - 1);
    name[cap - 1] = 0;
  }
  if (launches) *launches = s.launches;
  if (total_ms) *total_ms = s.total_ms;
  return SBDX_OK;
  SBDX_GUARD_END
}

int sbdx_work_counts(sbdx_handle_t h, double *counts) {
  SBDX_GUARD_BEGIN
  Context &c = ctx_of(h);
  if (!counts) throw Error(SBDX_ERR_ARG, "null counts");
  counts[0] = (double)c.A.ns;
  counts[1] = (double)c.A.nd;
  counts[2] = (double)c.B.ns;
  counts[3] = (double)c.B.nd;
  counts[4] = (double)c.A.ns * (double)c.B.ns;
  counts[5] = (double)c.N;
  return SBDX_OK;
  SBDX_GUARD_END
}

}  // extern "C"
