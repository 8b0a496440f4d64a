// Internal declarations of libsbdx (HIP, gfx950 only).
//
// Device-side conventions
//  * strings: one uint64 per string (norb <= 64), orbital k at bit k;
//  * vectors: alpha-major [ia*nb + ib] in the ALPHA-FIRST determinant basis
//    (alpha creators ascending, then beta creators ascending).  The API
//    converts from/to the reference's interleaved basis with the diagonal
//    sign S(Ia,Ib) = (-1)^#{(p in Ia, r in Ib): r < p}; in the alpha-first
//    basis the phase of an alpha excitation depends on the alpha string only
//    and that of a beta excitation on the beta string only;
//  * dense ERI: eri[pq*K + rs] = (pq|rs), pq = p*L + q, K = L*L.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace sbdx {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define SBDX_HIP(call)                                                            \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                         \
      throw ::sbdx::Error(2, std::string(#call) + ": " + hipGetErrorString(e_) +  \
                                 " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"); \
  } while (0)

// Excitation tables of one spin set.  Singles entry: ket | pqidx<<32 |
// (sign<0)<<63 with pqidx = p*L + q (p bra-only, q ket-only); s_val = sign*F,
// F = h_pq + sum_{k in bra} ((pq|kk) - (pk|kq)).  Doubles: ket, orbitals
// i | j<<8 | a<<16 | b<<24 (i<j bra-only, a<b ket-only), signed value.
struct SpinTables {
  int n = 0;
  uint64_t *str = nullptr;     // [n]
  uint64_t *s_off = nullptr;   // [n+1]
  uint64_t *s_ent = nullptr;   // [ns]
  double *s_val = nullptr;     // [ns]
  uint64_t *d_off = nullptr;   // [n+1]
  uint32_t *d_ket = nullptr;   // [nd]
  uint32_t *d_orb = nullptr;   // [nd]
  double *d_val = nullptr;     // [nd]
  double *e1 = nullptr;        // [n]: sum_p h_pp + sum_{p<q}((pp|qq)-(pq|qp)) of the string
  double *jsum = nullptr;      // [n][L]: sum_{p in str} (pp|rr)
  size_t ns = 0, nd = 0;
  size_t max_s = 0;            // longest singles row
  bool alias = false;          // shares arrays with the other spin
  // Singles again in sliced-ELL form for the mixed kernels: rows sorted by
  // singles count (descending) into slices of 64 (one wavefront); within a
  // slice entry k of lane l sits at sl_off[slice] + 64*k + l, packed as
  // ket | pq<<16 | (sign<0)<<31; padding entries carry pq = K (a zero slot).
  uint32_t *ell = nullptr;
  uint32_t *perm = nullptr;    // [nslice*64] row of each slot, 0xffffffff for padding
  uint64_t *sl_off = nullptr;  // [nslice]
  uint32_t *sl_len = nullptr;  // [nslice]
  int nslice = 0;
  bool has_ell = false;
  // The same entries packed for the staged mixed kernel k_mixed_ell (same
  // perm and slices): (K + ket) << 4 | pq << 19 | (sign<0) << 31, padding
  // ((K + n) << 4); slice lengths sl_len2 rounded up to a multiple of 4 and
  // each group of 4 positions stored lane-major (word u of lane l of group g
  // at sl_off2 + 256 g + 4 l + u); needs K <= 4096 and K + n < 2^15.
  uint32_t *ell2 = nullptr;
  uint64_t *sl_off2 = nullptr;  // [nslice]
  uint32_t *sl_len2 = nullptr;  // [nslice]
  bool has_ell2 = false;
  // The same entries for the signed-triangle kernel k_mixed_sw (same
  // slices, lengths and group layout as ell2): ket << 4 | widx << 21, where
  // widx = tri(pq) for a positive single and Kt + tri(pq) for a negative one
  // (tri(a*L+b) = max(a,b)(max(a,b)+1)/2 + min(a,b), Kt = L(L+1)/2): w >> 17
  // is the byte offset of the signed staged ERI value, w & 0x1fff0 that of
  // the ket value relative to the ket rows; padding (n << 4).  Built when
  // the ERI matrix is symmetric in its second index pair, Kt <= 1024 and
  // n < 8192.
  uint32_t *ell3 = nullptr;
  bool has_ell3 = false;
  // The singles grouped by pair index pq (for the RDM's mixed term):
  // pm_off[K+1]; pm_ent = bra | ket << 32 | (sign<0) << 63, rows ascending
  // within a pair.  The same grouping as a sliced ELL over the pair bins
  // (n < 32768): bins by entry count into slices of 64 lanes, entry k of
  // lane l of slice s at pb[pb_off[s] + 64 k + l] = bra | ket << 15 |
  // (sign<0) << 31 (padding: bra = ket = n), pb_len[s] entries per lane,
  // pb_perm[64 s + l] the lane's bin (0xffffffff past K).
  uint64_t *pm_off = nullptr;
  uint64_t *pm_ent = nullptr;
  uint32_t *pb = nullptr, *pb_len = nullptr, *pb_perm = nullptr;
  uint64_t *pb_off = nullptr;
  int pb_nslice = 0;
  // The same singles split by ket range into `nparts` parts of `part_h`
  // kets (ket stored relative to its part), same perm and slices; built on
  // demand when one ket row does not fit in LDS.
  int nparts = 0, part_h = 0;
  std::vector<uint32_t *> pell;
  std::vector<uint64_t *> psl_off;
  std::vector<uint32_t *> psl_len;
};

struct ProfSlot {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  long launches = 0;
  double total_ms = 0.0;
};

// FP32 copies of the sigma operands and work vectors (mixed precision).
struct F32Operands {
  bool ready = false;
  float *s_val_a = nullptr, *d_val_a = nullptr, *s_val_b = nullptr, *d_val_b = nullptr;
  float *eri = nullptr, *hii = nullptr;
  float *x = nullptr, *y = nullptr, *T1 = nullptr, *T2 = nullptr;
  float *xb = nullptr, *yb = nullptr;  // block products, kcap vectors each
  int kcap = 0;
};

// Learned preconditioner (mlp.hip): FP32 master weights, gradient, Adam
// moments, and the per-step history a solve records for training pairs.
struct MlpState {
  float *prm = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
  int adam_t = 0;
  bool valid = false;
  int record_cap = 0;              // >0: the thick solver records (R, W, E) per expansion
  std::vector<double *> rec_R, rec_W;
  std::vector<double> rec_E;
  int rec_n = 0;
  // preconditioner use statistics of the last solve
  long used_mlp = 0, used_diag = 0;
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  int L = 0, K = 0;
  size_t na = 0, nb = 0, N = 0;
  double I0 = 0.0;
  double *eri = nullptr;   // [K*K]
  // (pq|rs) == (pq|sr) bitwise for every pq, r, s (real orbitals): the
  // mixed kernels may stage each ERI row over the triangle r >= s only
  bool eri_rs_sym = false;
  int Kt = 0;                  // L(L+1)/2
  uint16_t *tri_rs = nullptr;  // [Kt]: rs = a*L + b of triangle element a(a+1)/2 + b, a >= b
  double *h1 = nullptr;    // [K] spatial one-body
  SpinTables A, B;
  double *hii = nullptr;   // [N]
  F32Operands f32;
  // Same-spin blocks as dense symmetric matrices for the GEMM form of the
  // same-spin terms (row-major n x n; the beta block aliases the alpha one
  // when the string sets are equal), FP64 and FP32, built on first use.
  double *Hd_a = nullptr, *Hd_b = nullptr;
  float *Hf_a = nullptr, *Hf_b = nullptr;
  void *blas = nullptr;  // rocblas_handle bound to `stream`
  // Same-spin form per precision [0 FP64, 1 FP32]: -1 not chosen yet,
  // 0 sparse row kernels, 1 dense GEMM; the timings that decided it (ms).
  int same_mode[2] = {-1, -1};
  float same_ms[2][2] = {{0.f, 0.f}, {0.f, 0.f}};
  // Sliced-ELL mixed kernel configuration per form (0 FP64, 1 FP32 general;
  // 2 FP64, 3 FP32 spin-flip symmetric): staging width (pairs per LDS round;
  // 0 not tuned, -1 none fits, else 1, 2 or 4), slots per thread, and the
  // tuned time (ms).
  // split-FP16x3 GEMM of the symmetric FP32 same-spin term: mode (-1 not
  // decided, 0 FP32 SGEMM, 1 split), the deciding timings (ms: SGEMM,
  // split), the FP16 operands (Hh: [Hh; Hl; Hh] of H_alpha stacked per
  // column, 3n x n; Xh: [Xh | Xh | Xl] of the current x, n x 3n; Hl, Xl
  // unused), per-block maxima, and the two inverse power-of-two scales
  // (x's, H's)
  int fp64_switch_step = -1;  // last thick solve: step from which products were FP64 (-1 never)
  int split_mode = -1;
  float split_ms[2] = {0.f, 0.f};
  void *Hh = nullptr, *Hl = nullptr, *Xh = nullptr, *Xl = nullptr;
  float *split_part = nullptr, *split_inv = nullptr;
  // Mixed alpha-beta term: 0 direct sliced-ELL kernels; 1 factored dense
  // (pq|rs) GEMM in the product's precision; 2 factored with the
  // split-FP16x3 GEMM for FP32 products.  Workspace of the factored form
  // (built on first use): D and G (fac_cap bytes each), x^T, the FP16
  // operands [Dh | Dh | Dl] and [Eh; El; Eh] (stacked per column) of the
  // split GEMM, and their
  // two inverse scales (D's, ERI's).
  int mixed_form = 0;
  void *fac_D = nullptr, *fac_G = nullptr, *fac_xT = nullptr;
  void *fac_D3 = nullptr, *fac_E3 = nullptr;
  float *fac_inv = nullptr;
  float *fac_part = nullptr;  // block maxima of |x| of the current factored FP32 product
  size_t fac_cap = 0;
  int mixed_eb[4] = {0, 0, 0, 0};
  int mixed_nt[4] = {0, 0, 0, 0};
  float mixed_best_ms[4] = {0.f, 0.f, 0.f, 0.f};
  // slots per thread of the row-parts mixed kernel, per form as mixed_eb
  // (0 until timed on first use)
  int parts_nt[4] = {0, 0, 0, 0};
  // the coarse start's sub-problem context (first coarse_sub_m strings of
  // each set), kept between solves of the same size; owned by this context
  Context *coarse_sub = nullptr;
  size_t coarse_sub_m = 0;
  // block products: 1 when the multi-vector kernel beat per-vector launches
  // on first use, 0 when not, -1 not decided
  int multi_mode[2] = {-1, -1};
  int precision = 0;       // 0: FP64 throughout; 1: FP32 sigma on corrections, FP64 restarts
  int verbose = 0;         // 1: one stderr line per Davidson step (E0, |r|)
  // spin-flip sectors of the thick-restart solves: 0 general form; 1 the
  // even sector (x = x^T), 2 the odd sector (x = -x^T), 3 both, the lower
  // result kept.  sym_par (+1 / -1) is the sector of the symmetric products
  // being computed.
  int symmetric = 0;
  int sym_par = 1;
  // results of the last two-sector solve: energies, products, converged
  // flags, abandoned (odd stage stopped once it lay above the even result),
  // the sector returned (1 even, 2 odd)
  double sector_E[2] = {0.0, 0.0};
  int sector_products[2] = {0, 0};
  int sector_state[2] = {0, 0};  // 0 not run, 1 converged, 2 not converged, 3 abandoned
  int sector_chosen = 0;
  // thick solves stop early (status 6) once E0 - |r| exceeds this value
  // (the odd stage of a two-sector solve); +inf: never
  double abandon_above = 1e300;
  // hipGraph replay of the FP32 symmetric sigma core (fixed FP32 staging
  // buffers): captured once after an eager warm-up when use_graphs is set
  int use_graphs = 0;
  bool capturing = false;  // no profiling events while a stream capture runs
  hipGraphExec_t g_sym32 = nullptr;
  int g_sym32_par = 1;     // the sector the graph was captured for
  // Coarse start of the thick-restart solves: the problem on the first
  // coarse_m strings of each set solved to coarse_eps first (0: off), and
  // what it cost in the last solve.
  size_t coarse_m = 0;
  double coarse_eps = 1e-2;
  int coarse_products = 0;
  double coarse_energy = 0.0, coarse_seconds = 0.0;
  // host copies of the unpacked strings and dense spatial integrals
  std::vector<uint64_t> h_astr, h_bstr;
  std::vector<double> h_h1, h_eri;
  int restart_keep = 0;    // 0: restart from W alone (reference); k: thick restart keeping k Ritz vectors
  std::vector<double *> spareY, spareHY;  // thick-restart staging vectors
  // Block Davidson: basis columns and products in an FP64 bank (blk_r64
  // columns) and an FP32 bank (blk_m - blk_r64 columns), residual block R
  // (blk_block FP64 vectors), reduction partials, small matrices; the basis
  // size asked for (0: as large as memory allows, up to 64) and what the
  // last solve used.
  double *blk_C64 = nullptr, *blk_HC64 = nullptr, *blk_R = nullptr, *blk_part = nullptr, *blk_small = nullptr;
  float *blk_C32 = nullptr, *blk_HC32 = nullptr;
  int blk_m = 0, blk_r64 = 0, blk_r64cap = 0, blk_mixed = -1, blk_block = 0, blk_m_alloc_req = -1;
  int blk_m_req = 0, blk_m_used = 0, blk_keep_used = 0, blk_restarts = 0;
  std::vector<double> blk_E, blk_resid;  // Ritz values and residual norms of the last block solve
  std::vector<int> blk_par;              // and the roots' spin-flip sectors (+1, -1; 0: sectors off)
  MlpState mlp;
  int precond = 0;         // 0: diagonal R/(E0-d); 1: MLP with diagonal fallback on stall
  // work vectors
  double *vin = nullptr, *vout = nullptr;   // API staging (alpha-first)
  double *T1 = nullptr, *T2 = nullptr;      // transpose scratch [N]
  double *part = nullptr;                   // reduction partials
  size_t part_cap = 0;
  // Davidson state
  int dav_nb = 0;
  std::vector<double *> C, HC;
  std::vector<float *> HCf;  // FP32 storage of FP32 products (thick solver, symmetric mixed mode)
  double *W = nullptr, *R = nullptr;
  double *dsmall = nullptr;                 // small device matrices/scalars
  double *hpin = nullptr;                   // pinned host staging for the subspace eigensolver
  std::vector<double> trace;
  bool have_vector = false;
  size_t lds_limit = 65536;
  int n_cu = 256;
  // profiling
  bool prof = false;
  std::map<std::string, ProfSlot> prof_slots;
  std::vector<std::string> prof_order;
  std::vector<void *> allocs;
};

// small-matrix device layout inside Context::dsmall (doubles)
// [0, nbmax^2)            H (subspace matrix, column-major, ld = nbmax)
// [nbmax^2, 2 nbmax^2)    U (eigenvectors)
// [2 nbmax^2, +nbmax)     E (eigenvalues ascending)
// then scalars: norms, residual
constexpr int kNbMax = 64;
constexpr size_t kSmallDoubles = 2 * kNbMax * kNbMax + kNbMax + 64 + kNbMax;
constexpr int kScalarOffset = 2 * kNbMax * kNbMax + kNbMax;
constexpr int kCoefOffset = kScalarOffset + 64;  // kNbMax projection coefficients

// ---------------------------------------------------------------- helpers
void *dalloc(Context &c, size_t bytes);
void dfree_all(Context &c);
void dfree(Context &c, void *p);  // one dalloc'ed buffer
Context *make_sub_context(const Context &c, size_t m);  // first m strings of each set
void destroy_context(Context *c);
void *blas_handle(Context &c);     // rocblas_handle bound to the stream

// Device scratch released at scope exit (callers synchronize first).
struct TmpBuf {
  void *p = nullptr;
  explicit TmpBuf(size_t bytes);
  ~TmpBuf();
  TmpBuf(const TmpBuf &) = delete;
  TmpBuf &operator=(const TmpBuf &) = delete;
};

struct ProfScope {
  Context &c;
  hipEvent_t b = nullptr, e = nullptr;
  const char *name;
  ProfScope(Context &cx, const char *nm);
  ~ProfScope();
};

void check_launch(const char *what);

// tables.hip
void build_spin_tables(Context &c, SpinTables &t, const std::vector<uint64_t> &strs);
void build_singles_ell(Context &c, SpinTables &t);
void build_singles_ell_parts(Context &c, SpinTables &t, int nparts);
void compute_diag(Context &c);
void sign_convert(Context &c, const double *in, double *out);      // out = S*in
void sign_convert_noprof(Context &c, const double *in, double *out);  // out = S*in, no profiling events
void sign_convert_add(Context &c, const double *in, double *out);  // out += S*in

// sigma.hip
void sigma_alpha_first(Context &c, const double *x, double *y);  // y = H' x (overwrite)
void sigma_alpha_first_f32(Context &c, const double *x, double *y);  // same, FP32 arithmetic
void ensure_f32(Context &c);
// y_v = H' x_v for k contiguous vectors (x + v*N), FP64 / FP32 arithmetic
void sigma_block(Context &c, int k, const double *x, double *y);
// spin-flip symmetric form (equal string sets, x = x^T as na x na): y = H' x
bool sigma_sym_available(Context &c);
void sigma_sym(Context &c, const double *x, double *y);
void sigma_sym_f32(Context &c, const double *x, double *y);
void sigma_sym_f32_to(Context &c, const double *x, float *y);
void sigma_sym_f32_raw(Context &c, const float *x, float *y);  // FP32 vectors in and out, no graph
double sym_defect(Context &c, const double *v, int p);  // max|v - p v^T| / max|v|, p = +-1
void symmetrize(Context &c, double *v, int p);         // v = (v + p v^T)/2
void sigma_block_f32(Context &c, int k, const double *x, double *y);
void sigma_block_f32_raw(Context &c, int k, const float *x, float *y);  // FP32 vectors in and out
void same_dense_release(Context &c);  // frees the rocBLAS handle
void graphs_release(Context &c);      // frees captured graphs
// eigenpairs of a symmetric n x n column-major A (overwritten), ascending:
// E[n], eigenvectors in the columns of Q (n x n)
void host_syev(int n, double *A, double *Q, double *E);
// H y = theta S y (S positive definite), n x n column-major with leading
// dimension ld: S-orthonormal eigenvectors Y (n x n, ld n), E ascending
void host_sygv(int n, const double *H, const double *S, int ld, double *Y, double *E);
void transpose_into(Context &c, const double *in, double *out, int R, int Cc);  // out = in^T

// davidson.hip
int davidson_solve(Context &c, int max_it, int nb, double eps, double max_time, double *energy,
                   int *n_steps);
int davidson_block(Context &c, int nroots, int block, int max_it, double eps, double max_time,
                   bool mixed, const double *W0, double *Wout, double *Eout, int *n_products);
int davidson_solve_thick(Context &c, int max_steps, int nb, int keep, double eps, double max_time,
                         bool mixed, double *energy, int *n_steps);
// the thick solve in the spin-flip sector(s) of c.symmetric (mode 3: both,
// the lower kept)
int davidson_solve_sectors(Context &c, int max_steps, int nb, int keep, double eps, double max_time,
                           bool mixed, double *energy, int *n_steps);
double device_dot(Context &c, const double *x, const double *y, size_t n);

// rdm.hip
void compute_rdm(Context &c, const double *Wp, double *one, double *two);
// the same with working memory from the static scratch array of rdm.hip
// (no allocation; errors when the instance needs more than it holds)
void compute_rdm_noalloc(Context &c, const double *Wp, double *one, double *two);

// mlp.hip
int mlp_nparam();
int mlp_nfeat();
void mlp_ensure(Context &c);
void mlp_set(Context &c, const float *host);
void mlp_get(Context &c, float *host);
void mlp_forward(Context &c, const unsigned short *dfeat, size_t rows, float *dout);
double mlp_gradient(Context &c, const unsigned short *dfeat, const float *dtarget, const uint32_t *didx,
                    size_t nbatch);
void mlp_adam_step(Context &c, float lr);
void mlp_precond(Context &c, const double *R, const double *E, double *Cout);
void mlp_features(Context &c, const double *R, double e, unsigned short *dout);
size_t mlp_collect(Context &c, size_t max_pairs, uint64_t seed, unsigned short *feat_out, float *tgt_out);
void mlp_train(Context &c, const unsigned short *dfeat, const float *dtgt, size_t n, int epochs, int batch,
               float lr, uint64_t seed, double *loss);

}  // namespace sbdx
