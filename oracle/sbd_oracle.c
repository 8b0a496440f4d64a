/*
 * sbd_oracle.c -- CPU restatement of the reference TPB (tensor-product basis)
 * Hamiltonian path: integrals, alpha/beta excitation tables, Slater-Condon
 * matrix elements, sigma = H*C, the diagonal, the single-root Davidson solver
 * and the 1-/2-RDM.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product path (libsbdx.so) never links or calls it.
 *
 * Conventions follow the reference exactly (interleaved spin orbitals: alpha
 * orbital k -> spin-orbital 2k, beta orbital k -> 2k+1; CI vector alpha-major
 * [ia*Nb + ib]; integrals in the reference's spin-orbital storage).
 * Strings are single 64-bit masks (norb <= 64), full determinants 128-bit.
 *
 * Reference citations are relative to the sbd-pub source tree.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------ */
/* bit helpers                                                         */
/* ------------------------------------------------------------------ */
static inline int popc128(u128 x) {
  return __builtin_popcountll((uint64_t)x) +
         __builtin_popcountll((uint64_t)(x >> 64));
}
static inline int bit128(u128 x, int i) { return (int)((x >> i) & 1); }

/* Interleave an alpha and a beta string into a full determinant.
 * Follows chemistry/basic/determinants.h:96-142 (alpha k -> 2k, beta k -> 2k+1). */
static u128 interleave(uint64_t a, uint64_t b, int L) {
  u128 d = 0;
  for (int k = 0; k < L; k++) {
    if ((a >> k) & 1) d |= (u128)1 << (2 * k);
    if ((b >> k) & 1) d |= (u128)1 << (2 * k + 1);
  }
  return d;
}

/* Fermionic sign factor of the reference, chemistry/basic/determinants.h:230-281:
 * count the occupied bits in [s, e), flip once per bit, and flip once more
 * when bit s itself is occupied. */
static double parity(u128 det, int s, int e) {
  u128 lo = ((u128)1 << s) - 1;
  u128 hi = ((u128)1 << e) - 1;
  int n = popc128(det & (hi ^ lo));
  double sg = (n & 1) ? -1.0 : 1.0;
  if (bit128(det, s)) sg = -sg;
  return sg;
}

/* ------------------------------------------------------------------ */
/* integrals (chemistry/basic/integrals.h:10-22, 64-107;               */
/*            chemistry/basic/makeintegrals.h:123-174)                 */
/* ------------------------------------------------------------------ */
typedef struct {
  int L;        /* spatial orbitals */
  double I0;
  double *I1;   /* (2L)x(2L) row-major, [i*2L + j] */
  double *I2;   /* compressed 8-fold spatial store */
  size_t nI2;
  double *Dir;  /* Direct(i,j)   at [i + 2L*j] */
  double *Exc;  /* Exchange(i,j) at [i + 2L*j] */
} ints_t;

static inline int tri(int a, int b) {
  int mx = a > b ? a : b, mn = a > b ? b : a;
  return mx * (mx + 1) / 2 + mn;
}

/* twoInt::Value: zero unless the spins of (i,j) and of (k,l) agree; then the
 * 8-fold compressed spatial index. */
static inline double V2(const ints_t *g, int i, int j, int k, int l) {
  if (!(((i & 1) == (j & 1)) && ((k & 1) == (l & 1)))) return 0.0;
  int ij = tri(i >> 1, j >> 1), kl = tri(k >> 1, l >> 1);
  return g->I2[tri(ij, kl)];
}
static inline double V1(const ints_t *g, int i, int j) { return g->I1[i * 2 * g->L + j]; }

size_t orc_i2_size(int L) {
  size_t B = (size_t)L * (L + 1) / 2;
  return B * (B + 1) / 2;
}

/* SetupIntegrals (makeintegrals.h:123-174): fcidump rows (value,i,j,k,l),
 * 1-based orbitals; (0,0,0,0) -> I0; (i,j,0,0) -> one-body (both spins,
 * both orders); otherwise the compressed two-body store.  Later rows
 * overwrite earlier ones, as in the reference. */
int orc_setup_integrals(int L, size_t n, const double *val, const int *ijkl,
                        double *I0, double *I1, double *I2) {
  size_t n2 = orc_i2_size(L);
  int L2 = 2 * L;
  memset(I1, 0, sizeof(double) * (size_t)L2 * L2);
  memset(I2, 0, sizeof(double) * n2);
  *I0 = 0.0;
  for (size_t r = 0; r < n; r++) {
    int i = ijkl[4 * r], j = ijkl[4 * r + 1], k = ijkl[4 * r + 2], l = ijkl[4 * r + 3];
    double v = val[r];
    if (i == 0 && j == 0 && k == 0 && l == 0) {
      *I0 = v;
    } else if (k == 0 && l == 0) {
      int a = 2 * (i - 1), b = 2 * (j - 1);
      I1[a * L2 + b] = v;
      I1[(a + 1) * L2 + b + 1] = v;
      I1[b * L2 + a] = v;
      I1[(b + 1) * L2 + a + 1] = v;
    } else {
      I2[tri(tri(i - 1, j - 1), tri(k - 1, l - 1))] = v;
    }
  }
  return 0;
}

static void ints_init(ints_t *g, int L, double I0, const double *I1, const double *I2) {
  g->L = L;
  g->I0 = I0;
  int L2 = 2 * L;
  g->nI2 = orc_i2_size(L);
  g->I1 = malloc(sizeof(double) * L2 * L2);
  memcpy(g->I1, I1, sizeof(double) * L2 * L2);
  g->I2 = malloc(sizeof(double) * g->nI2);
  memcpy(g->I2, I2, sizeof(double) * g->nI2);
  g->Dir = malloc(sizeof(double) * L2 * L2);
  g->Exc = malloc(sizeof(double) * L2 * L2);
  for (int i = 0; i < L2; i++)
    for (int j = 0; j < L2; j++) {
      g->Dir[i + L2 * j] = V2(g, i, i, j, j);
      g->Exc[i + L2 * j] = V2(g, i, j, j, i);
    }
}
static void ints_free(ints_t *g) {
  free(g->I1); free(g->I2); free(g->Dir); free(g->Exc);
}

/* ------------------------------------------------------------------ */
/* Slater-Condon elements (determinants.h:434-501)                     */
/* ------------------------------------------------------------------ */
/* ZeroExcite: diagonal element, including I0. */
static double zero_excite(const ints_t *g, u128 det) {
  int L2 = 2 * g->L;
  double e = 0.0;
  for (int i = 0; i < L2; i++) {
    if (!bit128(det, i)) continue;
    e += V1(g, i, i);
    for (int j = i + 1; j < L2; j++) {
      if (!bit128(det, j)) continue;
      e += g->Dir[i + L2 * j];
      if ((i & 1) == (j & 1)) e -= g->Exc[i + L2 * j];
    }
  }
  return e + g->I0;
}

/* OneExcite(det, i, a): det is the bra, i the bra-only spin orbital, a the
 * ket-only one.  Sum over every occupied spin orbital j of the bra, ascending. */
static double one_excite(const ints_t *g, u128 det, int i, int a) {
  double sg = parity(det, i < a ? i : a, i < a ? a : i);
  double e = V1(g, a, i);
  int L2 = 2 * g->L;
  for (int j = 0; j < L2; j++) {
    if (!bit128(det, j)) continue;
    e += (V2(g, a, i, j, j) - V2(g, a, j, j, i));
  }
  return e * sg;
}

/* TwoExcite(det, i, j, a, b): i,j bra-only, a,b ket-only. */
static double two_excite(const ints_t *g, u128 det, int i, int j, int a, int b) {
  int I = i < j ? i : j, J = i < j ? j : i;
  int A = a < b ? a : b, B = a < b ? b : a;
  double sg = parity(det, I < A ? I : A, I < A ? A : I);
  sg *= parity(det, J < B ? J : B, J < B ? B : J);
  if (A > J || B < I) sg = -sg;
  return sg * (V2(g, A, I, B, J) - V2(g, A, J, B, I));
}

/* exported primitives (for brute-force pinning in the tests) */
static u128 mk128(uint64_t lo, uint64_t hi) { return ((u128)hi << 64) | lo; }
double orc_parity(uint64_t lo, uint64_t hi, int s, int e) { return parity(mk128(lo, hi), s, e); }
void orc_interleave(uint64_t a, uint64_t b, int L, uint64_t *lo, uint64_t *hi) {
  u128 d = interleave(a, b, L);
  *lo = (uint64_t)d;
  *hi = (uint64_t)(d >> 64);
}

/* ------------------------------------------------------------------ */
/* excitation tables (chemistry/tpb/helper.h:185-323)                  */
/* For each bra string, every ket string of the set differing by one  */
/* (singles) or two (doubles) orbitals, in ascending ket index; cran   */
/* holds spin orbitals [cr..., an...] with cr = bra-only, an = ket-only*/
/* (2*orb for alpha, 2*orb+1 for beta).                               */
/* ------------------------------------------------------------------ */
typedef struct {
  size_t n;
  size_t *s_off, *d_off;  /* n+1 */
  uint32_t *s_ket, *d_ket;
  int *s_cran;            /* 2 per single */
  int *d_cran;            /* 4 per double */
} table_t;

static void build_table(table_t *t, const uint64_t *s, size_t n, int spin) {
  t->n = n;
  t->s_off = calloc(n + 1, sizeof(size_t));
  t->d_off = calloc(n + 1, sizeof(size_t));
  size_t *sc = calloc(n, sizeof(size_t)), *dc = calloc(n, sizeof(size_t));
#pragma omp parallel for schedule(dynamic, 16)
  for (size_t i = 0; i < n; i++) {
    size_t a = 0, b = 0;
    for (size_t j = 0; j < n; j++) {
      int d = __builtin_popcountll(s[i] ^ s[j]);
      if (d == 2) a++;
      else if (d == 4) b++;
    }
    sc[i] = a;
    dc[i] = b;
  }
  for (size_t i = 0; i < n; i++) {
    t->s_off[i + 1] = t->s_off[i] + sc[i];
    t->d_off[i + 1] = t->d_off[i] + dc[i];
  }
  free(sc); free(dc);
  t->s_ket = malloc(sizeof(uint32_t) * (t->s_off[n] + 1));
  t->d_ket = malloc(sizeof(uint32_t) * (t->d_off[n] + 1));
  t->s_cran = malloc(sizeof(int) * 2 * (t->s_off[n] + 1));
  t->d_cran = malloc(sizeof(int) * 4 * (t->d_off[n] + 1));
#pragma omp parallel for schedule(dynamic, 16)
  for (size_t i = 0; i < n; i++) {
    size_t ps = t->s_off[i], pd = t->d_off[i];
    for (size_t j = 0; j < n; j++) {
      uint64_t x = s[i] ^ s[j];
      int d = __builtin_popcountll(x);
      if (d != 2 && d != 4) continue;
      uint64_t cr = s[i] & ~s[j], an = s[j] & ~s[i];
      if (d == 2) {
        t->s_ket[ps] = (uint32_t)j;
        t->s_cran[2 * ps] = 2 * __builtin_ctzll(cr) + spin;
        t->s_cran[2 * ps + 1] = 2 * __builtin_ctzll(an) + spin;
        ps++;
      } else {
        int c0 = __builtin_ctzll(cr), c1 = 63 - __builtin_clzll(cr);
        int a0 = __builtin_ctzll(an), a1 = 63 - __builtin_clzll(an);
        t->d_ket[pd] = (uint32_t)j;
        t->d_cran[4 * pd] = 2 * c0 + spin;
        t->d_cran[4 * pd + 1] = 2 * c1 + spin;
        t->d_cran[4 * pd + 2] = 2 * a0 + spin;
        t->d_cran[4 * pd + 3] = 2 * a1 + spin;
        pd++;
      }
    }
  }
}
static void free_table(table_t *t) {
  free(t->s_off); free(t->d_off); free(t->s_ket); free(t->d_ket);
  free(t->s_cran); free(t->d_cran);
}

/* ------------------------------------------------------------------ */
/* context                                                            */
/* ------------------------------------------------------------------ */
typedef struct {
  ints_t g;
  int L;
  size_t na, nb;
  uint64_t *astr, *bstr;
  table_t ta, tb;
  int nthreads;
} orc_t;

void *orc_create(int L, double I0, const double *I1, const double *I2,
                 const uint64_t *astr, size_t na, const uint64_t *bstr, size_t nb,
                 int nthreads) {
  if (L < 1 || L > 64) return NULL;
  orc_t *o = calloc(1, sizeof(orc_t));
  o->L = L;
  o->na = na;
  o->nb = nb;
  o->nthreads = nthreads > 0 ? nthreads : omp_get_max_threads();
  omp_set_num_threads(o->nthreads);
  ints_init(&o->g, L, I0, I1, I2);
  o->astr = malloc(sizeof(uint64_t) * na);
  o->bstr = malloc(sizeof(uint64_t) * nb);
  memcpy(o->astr, astr, sizeof(uint64_t) * na);
  memcpy(o->bstr, bstr, sizeof(uint64_t) * nb);
  build_table(&o->ta, o->astr, na, 0);
  build_table(&o->tb, o->bstr, nb, 1);
  return o;
}

void orc_destroy(void *h) {
  orc_t *o = h;
  if (!o) return;
  ints_free(&o->g);
  free_table(&o->ta);
  free_table(&o->tb);
  free(o->astr); free(o->bstr);
  free(o);
}

/* table export: spin 0 = alpha, 1 = beta; kind 1 = singles, 2 = doubles.
 * With off == NULL only the entry count is returned. */
size_t orc_table(void *h, int spin, int kind, size_t *off, uint32_t *ket, int *cran) {
  orc_t *o = h;
  table_t *t = spin ? &o->tb : &o->ta;
  size_t n = t->n;
  size_t *src_off = kind == 1 ? t->s_off : t->d_off;
  size_t cnt = src_off[n];
  if (!off) return cnt;
  memcpy(off, src_off, sizeof(size_t) * (n + 1));
  memcpy(ket, kind == 1 ? t->s_ket : t->d_ket, sizeof(uint32_t) * cnt);
  memcpy(cran, kind == 1 ? t->s_cran : t->d_cran, sizeof(int) * cnt * (kind == 1 ? 2 : 4));
  return cnt;
}

/* diagonal, chemistry/tpb/qcham.h:289-340: hii = ZeroExcite(det). */
void orc_diag(void *h, double *hii) {
  orc_t *o = h;
  size_t nb = o->nb;
#pragma omp parallel for schedule(static)
  for (size_t ia = 0; ia < o->na; ia++)
    for (size_t ib = 0; ib < nb; ib++)
      hii[ia * nb + ib] = zero_excite(&o->g, interleave(o->astr[ia], o->bstr[ib], o->L));
}

/* ------------------------------------------------------------------ */
/* sigma, chemistry/tpb/mult.h:625-881 on one rank: Wb += hii*T, then   */
/* the three task types in the order alpha (S then D), beta (S then D), */
/* alpha-single x beta-single.  sigma is ACCUMULATED (caller zeroes).   */
/* Every element of a bra row receives its contributions in exactly that */
/* order, so one row at a time gives the same sums as the task loops.   */
/* ------------------------------------------------------------------ */
static void sigma_row(const orc_t *o, const double *hii, const double *T, double *W, size_t ia) {
  const size_t nb = o->nb;
  const ints_t *g = &o->g;
  const table_t *ta = &o->ta, *tb = &o->tb;
  for (size_t ib = 0; ib < nb; ib++) W[ia * nb + ib] += hii[ia * nb + ib] * T[ia * nb + ib];
  /* type 2: alpha singles and doubles */
  for (size_t ib = 0; ib < nb; ib++) {
    u128 D = interleave(o->astr[ia], o->bstr[ib], o->L);
    size_t bra = ia * nb + ib;
    for (size_t p = ta->s_off[ia]; p < ta->s_off[ia + 1]; p++) {
      double e = one_excite(g, D, ta->s_cran[2 * p], ta->s_cran[2 * p + 1]);
      W[bra] += e * T[(size_t)ta->s_ket[p] * nb + ib];
    }
    for (size_t p = ta->d_off[ia]; p < ta->d_off[ia + 1]; p++) {
      const int *c = ta->d_cran + 4 * p;
      double e = two_excite(g, D, c[0], c[1], c[2], c[3]);
      W[bra] += e * T[(size_t)ta->d_ket[p] * nb + ib];
    }
  }
  /* type 1: beta singles and doubles */
  for (size_t ib = 0; ib < nb; ib++) {
    u128 D = interleave(o->astr[ia], o->bstr[ib], o->L);
    size_t bra = ia * nb + ib;
    for (size_t p = tb->s_off[ib]; p < tb->s_off[ib + 1]; p++) {
      double e = one_excite(g, D, tb->s_cran[2 * p], tb->s_cran[2 * p + 1]);
      W[bra] += e * T[ia * nb + tb->s_ket[p]];
    }
    for (size_t p = tb->d_off[ib]; p < tb->d_off[ib + 1]; p++) {
      const int *c = tb->d_cran + 4 * p;
      double e = two_excite(g, D, c[0], c[1], c[2], c[3]);
      W[bra] += e * T[ia * nb + tb->d_ket[p]];
    }
  }
  /* type 0: alpha single x beta single */
  for (size_t ib = 0; ib < nb; ib++) {
    u128 D = interleave(o->astr[ia], o->bstr[ib], o->L);
    size_t bra = ia * nb + ib;
    for (size_t p = ta->s_off[ia]; p < ta->s_off[ia + 1]; p++) {
      size_t ja = ta->s_ket[p];
      int cra = ta->s_cran[2 * p], ana = ta->s_cran[2 * p + 1];
      for (size_t q = tb->s_off[ib]; q < tb->s_off[ib + 1]; q++) {
        double e = two_excite(g, D, cra, tb->s_cran[2 * q], ana, tb->s_cran[2 * q + 1]);
        W[bra] += e * T[ja * nb + tb->s_ket[q]];
      }
    }
  }
}

/* rows [row_lo, row_hi) only (bounded samples) */
void orc_sigma_rows(void *h, const double *hii, const double *T, double *W,
                    size_t row_lo, size_t row_hi) {
  orc_t *o = h;
  if (row_hi > o->na) row_hi = o->na;
#pragma omp parallel for schedule(dynamic, 1)
  for (size_t ia = row_lo; ia < row_hi; ia++) sigma_row(o, hii, T, W, ia);
}

void orc_sigma(void *h, const double *hii, const double *T, double *W) {
  orc_t *o = h;
  orc_sigma_rows(h, hii, T, W, 0, o->na);
}

/* sigma restricted to a list of distinct bra alpha rows (CPU timing samples) */
void orc_sigma_rowlist(void *h, const double *hii, const double *T, double *W,
                       const int64_t *rows, size_t nrows) {
  orc_t *o = h;
#pragma omp parallel for schedule(dynamic, 1)
  for (size_t k = 0; k < nrows; k++) sigma_row(o, hii, T, W, (size_t)rows[k]);
}

/* ------------------------------------------------------------------ */
/* vector helpers: Kahan-compensated dot per static thread chunk, then  */
/* chunk sums added in thread order (framework/dm_vector.h:22-86).      */
/* ------------------------------------------------------------------ */
static double kdot(const double *x, const double *y, size_t n, int nth) {
  double part[256];
  if (nth > 256) nth = 256;
  for (int t = 0; t < nth; t++) {
    size_t chunk = n / nth, rem = n % nth;
    size_t lo = t * chunk + (t < (int)rem ? t : rem);
    size_t hi = lo + chunk + (t < (int)rem ? 1 : 0);
    double s = 0.0, c = 0.0;
    for (size_t i = lo; i < hi; i++) {
      double v = x[i] * y[i] - c;
      double tt = s + v;
      c = (tt - s) - v;
      s = tt;
    }
    part[t] = s;
  }
  double r = 0.0;
  for (int t = 0; t < nth; t++) r += part[t];
  return r;
}
static double knormalize(double *x, size_t n, int nth) {
  double r = sqrt(kdot(x, x, n, nth));
  double f = 1.0 / r;
  for (size_t i = 0; i < n; i++) x[i] *= f;
  return r;
}

/* Symmetric eigensolver for the Rayleigh-Ritz step (stands in for LAPACK
 * dsyev, framework/hp_numeric.h:503-522: eigenvalues ascending, eigenvectors
 * in columns, column-major with leading dimension lda).  Cyclic Jacobi
 * rotations followed by an ascending sort. */
int orc_syev(int n, double *A, int lda, double *w) {
  double *V = calloc((size_t)n * n, sizeof(double));
  for (int i = 0; i < n; i++) V[i + i * n] = 1.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        double a = A[i + j * lda];
        tot += a * a;
        if (i != j) off += a * a;
      }
    if (off <= 1e-32 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        double apq = A[p + q * lda];
        if (apq == 0.0) continue;
        double app = A[p + p * lda], aqq = A[q + q * lda];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; k++) {
          double akp = A[k + p * lda], akq = A[k + q * lda];
          A[k + p * lda] = c * akp - s * akq;
          A[k + q * lda] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          double apk = A[p + k * lda], aqk = A[q + k * lda];
          A[p + k * lda] = c * apk - s * aqk;
          A[q + k * lda] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          double vkp = V[k + p * n], vkq = V[k + q * n];
          V[k + p * n] = c * vkp - s * vkq;
          V[k + q * n] = s * vkp + c * vkq;
        }
      }
  }
  int *idx = malloc(sizeof(int) * n);
  for (int i = 0; i < n; i++) { idx[i] = i; w[i] = A[i + i * lda]; }
  for (int i = 1; i < n; i++) {  /* insertion sort of eigenvalues */
    int k = idx[i];
    int j = i - 1;
    while (j >= 0 && w[idx[j]] > w[k]) { idx[j + 1] = idx[j]; j--; }
    idx[j + 1] = k;
  }
  double *ws = malloc(sizeof(double) * n);
  for (int i = 0; i < n; i++) ws[i] = w[idx[i]];
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) A[i + j * lda] = V[i + idx[j] * n];
  memcpy(w, ws, sizeof(double) * n);
  free(ws); free(idx); free(V);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Davidson, chemistry/tpb/davidson.h:667-930 (one rank): nb-vector     */
/* restarts, diagonal preconditioner R/(E0-d) with a 1e-12 guard,       */
/* modified Gram-Schmidt, Rayleigh-Ritz, stop when |E0 W - H W| < eps. */
/* trace (optional, 2 doubles per step: E0, |r|) records the history.  */
/* ------------------------------------------------------------------ */
int orc_davidson(void *h, const double *hii, double *W, int max_it, int nb_blk,
                 double eps, double *E_out, int *steps_out, double *trace, int trace_cap) {
  orc_t *o = h;
  size_t N = o->na * o->nb;
  int nth = o->nthreads;
  const double eps_reg = 1.0e-12;
  double **C = malloc(sizeof(double *) * nb_blk);
  double **HC = malloc(sizeof(double *) * nb_blk);
  for (int k = 0; k < nb_blk; k++) {
    C[k] = malloc(sizeof(double) * N);
    HC[k] = malloc(sizeof(double) * N);
  }
  double *R = malloc(sizeof(double) * N);
  double *H = calloc((size_t)nb_blk * nb_blk, sizeof(double));
  double *U = calloc((size_t)nb_blk * nb_blk, sizeof(double));
  double *E = calloc(nb_blk, sizeof(double));
  int steps = 0, done = 0;
  double E0 = 0.0;
  for (int it = 0; it < max_it && !done; it++) {
    memcpy(C[0], W, sizeof(double) * N);
    for (int ib = 0; ib < nb_blk; ib++) {
      memset(HC[ib], 0, sizeof(double) * N);
      orc_sigma(o, hii, C[ib], HC[ib]);
      for (int jb = 0; jb <= ib; jb++) {
        H[jb + nb_blk * ib] = kdot(C[jb], HC[ib], N, nth);
        H[ib + nb_blk * jb] = H[jb + nb_blk * ib];
      }
      for (int jb = 0; jb <= ib; jb++)
        for (int kb = 0; kb <= ib; kb++) U[jb + nb_blk * kb] = H[jb + nb_blk * kb];
      orc_syev(ib + 1, U, nb_blk, E);
      E0 = E[0];
#pragma omp parallel for schedule(static)
      for (size_t i = 0; i < N; i++) {
        double w = C[0][i] * U[0];
        double r = HC[0][i] * (-U[0]);
        for (int kb = 1; kb <= ib; kb++) {
          w += C[kb][i] * U[kb];
          r += HC[kb][i] * (-U[kb]);
        }
        W[i] = w;
        R[i] = r + E0 * w;
      }
      knormalize(W, N, nth);
      double nr = knormalize(R, N, nth);
      if (trace && steps < trace_cap) {
        trace[2 * steps] = E0;
        trace[2 * steps + 1] = nr;
      }
      steps++;
      if (nr < eps) { done = 1; break; }
      if (ib < nb_blk - 1) {
        double *c = C[ib + 1];
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < N; i++) {
          double den = E0 - hii[i];
          c[i] = fabs(den) > eps_reg ? R[i] / den : R[i] / (den - eps_reg);
        }
        for (int kb = 0; kb < ib + 1; kb++) {
          double ol = -kdot(C[kb], c, N, nth);
          for (size_t i = 0; i < N; i++) c[i] += C[kb][i] * ol;
        }
        knormalize(c, N, nth);
      }
    }
  }
  for (int k = 0; k < nb_blk; k++) { free(C[k]); free(HC[k]); }
  free(C); free(HC); free(R); free(H); free(U); free(E);
  if (E_out) *E_out = E0;
  if (steps_out) *steps_out = steps;
  return done ? 0 : 1;
}

/* ------------------------------------------------------------------ */
/* RDM, chemistry/tpb/correlation.h:378-662 with the contributions of   */
/* chemistry/basic/correlation.h:16-229.                                */
/* one[s*L*L + i + L*j], two[(s+2t)*L^4 + i + L j + L^2 k + L^3 l].      */
/* ------------------------------------------------------------------ */
static void zero_diff(u128 D, double w, int L, double *one, double *two) {
  size_t L2 = (size_t)L * L, L4 = L2 * L2;
  int occ[128], n = 0;
  for (int i = 0; i < 2 * L; i++) if (bit128(D, i)) occ[n++] = i;
  double ww = w * w;
  for (int a = 0; a < n; a++) {
    int oi = occ[a] >> 1, si = occ[a] & 1;
    one[si * L2 + oi + (size_t)L * oi] += ww;
    for (int b = a + 1; b < n; b++) {
      int oj = occ[b] >> 1, sj = occ[b] & 1;
      two[(si + 2 * sj) * L4 + oi + (size_t)L * oj + L2 * oi + L2 * L * oj] += ww;
      two[(sj + 2 * si) * L4 + oj + (size_t)L * oi + L2 * oj + L2 * L * oi] += ww;
      if (si == sj) {
        two[(si + 2 * sj) * L4 + oi + (size_t)L * oj + L2 * oj + L2 * L * oi] -= ww;
        two[(sj + 2 * si) * L4 + oj + (size_t)L * oi + L2 * oi + L2 * L * oj] -= ww;
      }
    }
  }
}

static void one_diff(u128 D, double wi, double wj, int L, int i, int a, double *one, double *two) {
  size_t L2 = (size_t)L * L, L4 = L2 * L2;
  double sg = parity(D, i < a ? i : a, i < a ? a : i);
  int oi = i >> 1, si = i & 1, oa = a >> 1;
  double v = wi * wj * sg;
  one[si * L2 + oi + (size_t)L * oa] += v;
  for (int soj = 0; soj < 2 * L; soj++) {
    if (!bit128(D, soj)) continue;
    int oj = soj >> 1, sj = soj & 1;
    two[(si + 2 * sj) * L4 + oa + (size_t)oj * L + (size_t)oi * L2 + (size_t)oj * L2 * L] += v;
    two[(sj + 2 * si) * L4 + oj + (size_t)oa * L + (size_t)oj * L2 + (size_t)oi * L2 * L] += v;
    if (si == sj) {
      two[(si + 2 * sj) * L4 + oa + (size_t)oj * L + (size_t)oj * L2 + (size_t)oi * L2 * L] -= v;
      two[(sj + 2 * si) * L4 + oj + (size_t)oa * L + (size_t)oi * L2 + (size_t)oj * L2 * L] -= v;
    }
  }
}

static void two_diff(u128 D, double wi, double wj, int L, int i, int j, int a, int b, double *two) {
  size_t L2 = (size_t)L * L, L4 = L2 * L2;
  int I = i < j ? i : j, J = i < j ? j : i;
  int A = a < b ? a : b, B = a < b ? b : a;
  double sg = parity(D, I < A ? I : A, I < A ? A : I);
  sg *= parity(D, J < B ? J : B, J < B ? B : J);
  if (A > J || B < I) sg = -sg;
  int oi = I >> 1, si = I & 1, oa = A >> 1, sa = A & 1;
  int oj = J >> 1, sj = J & 1, ob = B >> 1, sb = B & 1;
  (void)sb;
  double v = sg * wi * wj;
  if (si == sa) {
    two[(si + 2 * sj) * L4 + oa + (size_t)L * ob + L2 * (oi + (size_t)L * oj)] += v;
    two[(sj + 2 * si) * L4 + ob + (size_t)L * oa + L2 * (oj + (size_t)L * oi)] += v;
  }
  if (si == (B & 1)) {
    two[(si + 2 * sj) * L4 + oa + (size_t)L * ob + L2 * (oj + (size_t)L * oi)] -= v;
    two[(sj + 2 * si) * L4 + ob + (size_t)L * oa + L2 * (oi + (size_t)L * oj)] -= v;
  }
}

void orc_rdm(void *h, const double *W, double *one, double *two) {
  orc_t *o = h;
  int L = o->L;
  size_t nb = o->nb;
  size_t n1 = 2 * (size_t)L * L, n2 = 4 * (size_t)L * L * L * L;
  const table_t *ta = &o->ta, *tb = &o->tb;
  int nth = o->nthreads;
  double *one_t = calloc(n1 * nth, sizeof(double));
  double *two_t = calloc(n2 * nth, sizeof(double));
#pragma omp parallel num_threads(nth)
  {
    int tid = omp_get_thread_num();
    double *o1 = one_t + n1 * tid, *o2 = two_t + n2 * tid;
#pragma omp for schedule(dynamic, 1)
    for (size_t ia = 0; ia < o->na; ia++) {
      for (size_t ib = 0; ib < nb; ib++) {
        u128 D = interleave(o->astr[ia], o->bstr[ib], L);
        size_t bra = ia * nb + ib;
        double wi = W[bra];
        zero_diff(D, wi, L, o1, o2);
        for (size_t p = ta->s_off[ia]; p < ta->s_off[ia + 1]; p++)
          one_diff(D, wi, W[(size_t)ta->s_ket[p] * nb + ib], L, ta->s_cran[2 * p],
                   ta->s_cran[2 * p + 1], o1, o2);
        for (size_t p = ta->d_off[ia]; p < ta->d_off[ia + 1]; p++) {
          const int *c = ta->d_cran + 4 * p;
          two_diff(D, wi, W[(size_t)ta->d_ket[p] * nb + ib], L, c[0], c[1], c[2], c[3], o2);
        }
        for (size_t p = tb->s_off[ib]; p < tb->s_off[ib + 1]; p++)
          one_diff(D, wi, W[ia * nb + tb->s_ket[p]], L, tb->s_cran[2 * p],
                   tb->s_cran[2 * p + 1], o1, o2);
        for (size_t p = tb->d_off[ib]; p < tb->d_off[ib + 1]; p++) {
          const int *c = tb->d_cran + 4 * p;
          two_diff(D, wi, W[ia * nb + tb->d_ket[p]], L, c[0], c[1], c[2], c[3], o2);
        }
        for (size_t p = ta->s_off[ia]; p < ta->s_off[ia + 1]; p++) {
          size_t ja = ta->s_ket[p];
          int cra = ta->s_cran[2 * p], ana = ta->s_cran[2 * p + 1];
          for (size_t q = tb->s_off[ib]; q < tb->s_off[ib + 1]; q++)
            two_diff(D, wi, W[ja * nb + tb->s_ket[q]], L, cra, tb->s_cran[2 * q], ana,
                     tb->s_cran[2 * q + 1], o2);
        }
      }
    }
  }
  memset(one, 0, sizeof(double) * n1);
  memset(two, 0, sizeof(double) * n2);
  for (int t = 0; t < nth; t++) {
    for (size_t i = 0; i < n1; i++) one[i] += one_t[n1 * t + i];
    for (size_t i = 0; i < n2; i++) two[i] += two_t[n2 * t + i];
  }
  free(one_t);
  free(two_t);
}

/* single matrix-element access for pinning tests: <bra|H|ket> with the
 * reference's dispatch (determinants.h:503-623: Hij on DetA=bra, DetB=ket
 * passes the KET as the det argument).  The TPB sigma instead evaluates on
 * the bra; both are returned so the tests can check they agree. */
double orc_hij_bra(void *h, uint64_t a_bra, uint64_t b_bra, uint64_t a_ket, uint64_t b_ket) {
  orc_t *o = h;
  int L = o->L;
  u128 Db = interleave(a_bra, b_bra, L), Dk = interleave(a_ket, b_ket, L);
  u128 cr = Db & ~Dk, an = Dk & ~Db;
  int nc = popc128(cr);
  if (nc == 0) return zero_excite(&o->g, Db);
  if (nc > 2) return 0.0;
  int c[2], d[2], k = 0, m = 0;
  for (int i = 0; i < 2 * L; i++) {
    if (bit128(cr, i)) c[k++] = i;
    if (bit128(an, i)) d[m++] = i;
  }
  if (nc == 1) return one_excite(&o->g, Db, c[0], d[0]);
  if (nc == 2) return two_excite(&o->g, Db, c[0], c[1], d[0], d[1]);
  return 0.0;
}
