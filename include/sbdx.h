/*
 * sbdx.h -- C ABI of the MI355X (gfx950) implementation of the sbd-pub
 * tensor-product-basis (TPB) Hamiltonian path: excitation tables, the
 * diagonal, sigma = H*C, the single-root Davidson solver and the 1-/2-RDM.
 *
 * Every entry point returns an int status (SBDX_OK == 0) and never throws or
 * exits; sbdx_last_error() gives the message of the last failure on the
 * calling thread.  Caller buffers are plain host pointers unless the name
 * ends in _device; they are never retained past the call.  Device memory is
 * owned by the opaque handle.  One host thread per handle.
 *
 * Vectors are alpha-major, index ia*n_bdet + ib, in the reference's
 * determinant convention (interleaved spin orbitals, chemistry/basic/
 * determinants.h:96-142), so they are interchangeable with the reference's
 * own vectors.  Citations are relative to the sbd-pub source tree.
 */
#ifndef SBDX_H
#define SBDX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SBDX_OK 0
#define SBDX_ERR_ARG 1           /* invalid argument (sizes, pointers)       */
#define SBDX_ERR_HIP 2           /* HIP runtime or kernel launch failure     */
#define SBDX_ERR_UNSUPPORTED 3   /* instance outside supported limits        */
#define SBDX_ERR_ALLOC 4         /* device allocation failed                 */
#define SBDX_ERR_NOT_CONVERGED 5 /* solver stopped before reaching eps       */
#define SBDX_ERR_IO 6            /* file could not be read or parsed         */

typedef struct sbdx_context *sbdx_handle_t;

/* Message of the last failure on this thread ("" if none). */
const char *sbdx_last_error(void);
/* Library version, major*10000 + minor*100 + patch. */
int sbdx_version(void);

/*
 * Build an operator on the current HIP device.
 * Replaces MultTPBThrust::Init (chemistry/tpb/mult_thrust.h:59-94, 98-277)
 * together with MakeHelpers/RemakeHelpers (chemistry/tpb/helper.h:710-854,
 * 916-1226) for one rank: the excitation tables are built on the GPU.
 *   norb          spatial orbitals (1..64)
 *   bit_length    bits per size_t word of the string arrays (reference
 *                 default 20, chemistry/tpb/sbdiag.h:41)
 *   adet_words    n_adet strings x ceil(norb/bit_length) words, sorted and
 *                 unique as LoadAlphaDets leaves them
 *                 (chemistry/basic/makedeterminants.h:160-185)
 *   I0, I1, I2    core energy, oneInt::store ((2L)^2, row-major) and
 *                 twoInt::store (8-fold compressed, n_I2 = B(B+1)/2 with
 *                 B = L(L+1)/2), exactly as SetupIntegrals fills them
 *                 (chemistry/basic/makeintegrals.h:123-174)
 */
int sbdx_create(sbdx_handle_t *out, int norb, size_t bit_length,
                const uint64_t *adet_words, size_t n_adet,
                const uint64_t *bdet_words, size_t n_bdet,
                double I0, const double *I1, const double *I2, size_t n_I2);
int sbdx_destroy(sbdx_handle_t h);

/* n_adet, n_bdet and n_det = n_adet*n_bdet of the handle. */
int sbdx_dims(sbdx_handle_t h, size_t *n_adet, size_t *n_bdet, size_t *n_det);

/* Diagonal hii = ZeroExcite(det) including I0.
 * Replaces makeQChamDiagTerms (chemistry/tpb/qcham.h:289-340;
 * MultTPBThrust::makeQChamDiagTerms, mult_thrust.h:1585-1664). */
int sbdx_diag(sbdx_handle_t h, double *hii);

/* sigma: Wb (+)= H*Wk.  With accumulate != 0 the product is added to Wb, as
 * MultBase::run does (chemistry/basic/mult.h:19-103; CPU sbd::mult,
 * chemistry/tpb/mult.h:625-881); with accumulate == 0 Wb is overwritten.
 * The diagonal term uses the handle's own hii. */
int sbdx_sigma(sbdx_handle_t h, const double *Wk, double *Wb, int accumulate);
/* k products at once (no counterpart in the reference, whose operator takes
 * one vector per run): Wb[v] (+)= H*Wk[v] for v < k, host arrays of k
 * consecutive n_det vectors, 1 <= k <= 64.  The same-spin terms run as
 * batched GEMMs and the mixed term as one multi-vector kernel per group of
 * up to four vectors, so the table and ERI-row traffic is shared. */
int sbdx_sigma_block(sbdx_handle_t h, int k, const double *Wk, double *Wb, int accumulate);
/* The same k products in FP32 arithmetic (the block solver's correction
 * products in mixed precision); Wb is overwritten. */
int sbdx_sigma_block_fp32(sbdx_handle_t h, int k, const double *Wk, double *Wb);
/* Same with device pointers on the handle's stream (synchronous on return). */
int sbdx_sigma_device(sbdx_handle_t h, const double *d_Wk, double *d_Wb, int accumulate);

/*
 * Single-root Davidson, chemistry/tpb/davidson.h:667-930 and
 * chemistry/basic/davidson_thrust.h:85-373: restarts every nb vectors,
 * preconditioner R/(E0-d) with the 1e-12 guard, modified Gram-Schmidt,
 * Rayleigh-Ritz (subspace matrix built on the device, its eigenproblem of
 * at most 64 x 64 solved on the host as the reference's dsyev call is),
 * stop when |E0 W - H W| < eps, or when the
 * predicted end of the next step passes max_time seconds.
 * W (host, n_det) holds the initial vector on entry and the normalized
 * solution on exit.  *energy receives the lowest Ritz value of the last
 * step and *n_steps the number of sigma steps.  Returns
 * SBDX_ERR_NOT_CONVERGED (with valid outputs) if eps was not reached.
 */
int sbdx_davidson(sbdx_handle_t h, double *W, int max_it, int nb, double eps,
                  double max_time, double *energy, int *n_steps);

/*
 * Block Davidson for the lowest nroots eigenpairs (no counterpart in the
 * reference, whose Davidson is single-root): `block` (nroots <= block <= 16)
 * correction vectors per iteration go through one block sigma; converged
 * roots are locked (no corrections); thick restart on the lowest Ritz
 * vectors when the basis (sbdx_set_block_basis) is full.  Convergence when
 * |E_j W_j - H W_j| < eps for every j < nroots on a basis whose columns and
 * products are all FP64.  With SBDX_PREC_MIXED the correction vectors and
 * their products are FP32 in arithmetic and in storage (dot products and
 * combinations in FP64, Rayleigh-Ritz with the overlap matrix of the stored
 * columns), the roots are FP64 (with exact FP64 products at every restart
 * once max|r| < max(100 eps, 6e-5 |E|)), and convergence is confirmed on a
 * restart that
 * keeps the roots alone, with exact products.
 * The preconditioner is the handle's (sbdx_set_preconditioner).
 * W: host, nroots consecutive n_det vectors, linearly independent start
 * vectors on entry (block - nroots more are unit vectors on the lowest
 * diagonal elements), the Ritz vectors on exit; energies[nroots] ascending;
 * *n_products counts single-vector products.  max_it bounds iterations
 * (block expansions and restarts).  Returns SBDX_ERR_NOT_CONVERGED with
 * valid outputs when eps was not reached.
 */
int sbdx_davidson_block(sbdx_handle_t h, int nroots, int block, double *W, int max_it, double eps,
                        double max_time, double *energies, int *n_products);

/* Basis size of sbdx_davidson_block (no counterpart in the reference): m
 * columns (nroots + 2*block <= m <= 64), or 0 (default) for the largest m
 * up to 64, in steps of `block`, whose basis fits in 80% of free device
 * memory.  A thick restart keeps max(nroots + block, m/3) Ritz vectors
 * (at most m - block and 32). */
int sbdx_set_block_basis(sbdx_handle_t h, int m);

/* What the last sbdx_davidson_block used: basis size, Ritz vectors kept at
 * a restart, restarts, and per root (nroots entries, pointers may be null)
 * the final residual norm and the spin-flip sector (+1 even, -1 odd, 0
 * when sectors are off).  No counterpart in the reference. */
int sbdx_block_stats(sbdx_handle_t h, int *basis, int *keep, int *restarts, double *resid, int *parity,
                     int nroots);

/* Split form of sbdx_davidson for callers that keep the vector resident:
 * upload the start vector, solve on the device, download the solution. */
int sbdx_set_vector(sbdx_handle_t h, const double *W);
int sbdx_solve(sbdx_handle_t h, int max_it, int nb, double eps, double max_time,
               double *energy, int *n_steps);
int sbdx_get_vector(sbdx_handle_t h, double *W);

/* Arithmetic of the Davidson solves of this handle (no counterpart in the
 * reference, which is FP64 throughout):
 *   SBDX_PREC_FP64   every sigma in FP64 -- the reference's algorithm;
 *   SBDX_PREC_MIXED  the restart vector's sigma in FP64 (exact FP64
 *                    residual every cycle, convergence judged only on it),
 *                    the correction vectors' sigma in FP32; Gram-Schmidt,
 *                    Rayleigh-Ritz, residual and energy stay FP64. */
#define SBDX_PREC_FP64 0
#define SBDX_PREC_MIXED 1
int sbdx_set_precision(sbdx_handle_t h, int mode);

/* Restart policy of later solves (no counterpart in the reference):
 *   keep == 0   restart from W alone every nb steps (davidson.h:918-922);
 *   keep 1..8   thick restart: at each restart W gets an exact FP64 product
 *               and the exact residual decides convergence; the basis keeps
 *               W and the next keep-1 Ritz vectors (products by linear
 *               combination).  max_it then bounds max_it*nb products. */
int sbdx_set_restart(sbdx_handle_t h, int keep);

/* Form of the same-spin terms chosen for precision 0 (FP64) or 1 (FP32) (no
 * counterpart in the reference): -1 not chosen yet (no sigma of that
 * precision so far), 0 sparse row kernels, 1 dense GEMMs over the same-spin
 * blocks.  The choice is made on the first sigma of each precision by timing
 * both forms; ms_* receive those timings (0 when a form was not timed, e.g.
 * when SBDX_SAME_SPIN=sparse|dense in the environment fixed it). */
int sbdx_same_spin_form(sbdx_handle_t h, int precision, int *form, float *ms_sparse,
                        float *ms_dense);

/* Configuration of the mixed-term kernel chosen for `precision` 0 (FP64) or
 * 1 (FP32) of the general sigma, or 2 (FP64) or 3 (FP32) of the spin-flip
 * symmetric sigma, on the first product of that form by timing every
 * staging width (alpha singles staged in LDS per round: 1, 2, 4, or 16 for
 * the signed-triangle kernel, whose LDS elements are 16-byte pairs: 4 FP32
 * or 2 FP64 values, and whose staged ERI rows hold the r >= s triangle with
 * both signs) and slots per thread (1, 2, 4, 8; a row's slots are split
 * over blocks accordingly) that fit (no counterpart in the reference):
 * *width 0 before that, -1 when no width fits in LDS (row parts or the
 * CSR/global variants run), else 1, 2, 4 or 16; *slots_per_thread, *ms the
 * winner and its time (with width -1: the row-parts kernel's slots per
 * thread, timed the same way on its first use; SBDX_PARTS_NT fixes it);
 * *multi (forms 0
 * and 1) whether block products use the multi-vector kernel (1) or
 * per-vector launches (0), decided by timing on the first block product
 * (-1 before, and always -1 for forms 2 and 3).
 * SBDX_MIXED_EB=1|2|4|16 and SBDX_MIXED_NT=1|2|4|8 fix the coordinates. */
int sbdx_mixed_staging(sbdx_handle_t h, int precision, int *width, int *slots_per_thread,
                       float *ms, int *multi);

/* Mixed-precision thick solves (no counterpart in the reference): FP32
 * correction products cannot take the residual below roughly 1e-7 |H|, and
 * below that floor their errors feed back through the kept Ritz vectors.
 * When an exact restart residual fails to halve while below
 * 1000 x 2^-24 x |E0|, the solve continues with FP64 products (that restart
 * keeps W alone). *step receives the step of the last solve at which this
 * happened, or -1 when every correction product of it was FP32. */
int sbdx_fp64_switch_step(sbdx_handle_t h, int *step);

/* Form of the GEMM in the spin-flip symmetric FP32 same-spin term (no
 * counterpart in the reference): *mode -1 before the first such product, 0
 * FP32 SGEMM, 1 split-FP16x3 (x and H_alpha each written as 2^-e (hi + lo)
 * in FP16 with a power-of-two scale, about 22 significant bits; the three
 * FP16 products as one MFMA GEMM over k = 3n with FP32 accumulation, the
 * lo x lo product dropped). Decided
 * on the first product by timing both; *ms_sgemm, *ms_split those timings
 * (0 when SBDX_SPLIT_GEMM=0|1 in the environment fixed the mode). */
int sbdx_split_gemm(sbdx_handle_t h, int *mode, float *ms_sgemm, float *ms_split);

/* Form of the mixed alpha-beta term in every later product (no counterpart
 * in the reference, which evaluates each element directly,
 * chemistry/tpb/mult.h:625-881): 0 (default) the direct sliced-ELL kernels;
 * 1 the factored form, a dense (pq|rs) contraction on the matrix cores:
 * per batch of beta rows D[rs][(r,J)] = sum over the extended beta singles
 * of r (singles plus the occupied (kk) identity entries) of s_b x[J,j],
 * G = (pq|rs) D as one DGEMM or SGEMM, y[I,r] += the alpha side over G
 * minus the double-counted diagonal Coulomb term; 2 the same with the
 * split-FP16x3 GEMM for FP32 products (FP64 products use the DGEMM).  The
 * factored form costs 2 K^2 na nb flops per product (K = norb^2) and holds
 * up to 2 GiB of D and G plus x^T; products in it use the general form
 * even when the symmetric form is set, and no graph is captured. */
int sbdx_set_mixed_form(sbdx_handle_t h, int form);

/* The subspace eigensolver of the Davidson solves, on the host: eigenvalues
 * E[n] ascending and orthonormal eigenvectors (columns of V, n x n
 * column-major) of the symmetric n x n column-major A (not modified), by
 * Householder tridiagonalization and implicit QR with Wilkinson shifts.
 * Replaces the reference's LAPACK dsyev call (hp_numeric::MatHeev,
 * framework/hp_numeric.h:503-522); needs no handle or GPU. */
int sbdx_symmetric_eig(int n, const double *A, double *V, double *E);

/* The block solver's Rayleigh-Ritz step, on the host: H y = E S y for
 * symmetric H and positive definite S (n x n column-major, not modified),
 * by Cholesky S = L L^T and sbdx_symmetric_eig of L^-1 H L^-T; E[n]
 * ascending, S-orthonormal Y (columns, n x n column-major).  No
 * counterpart in the reference (its Davidson keeps an orthonormal basis);
 * needs no handle or GPU.  SBDX_ERR_UNSUPPORTED (3) when S is not
 * numerically positive definite. */
int sbdx_symmetric_geig(int n, const double *H, const double *S, double *Y, double *E);

/* Spin-flip symmetry (no counterpart in the reference): with equal alpha and
 * beta string sets, H commutes with the exchange of the two strings, so its
 * eigenvectors split into two sectors.  In the alpha-first basis of the
 * device a vector of the even sector is a symmetric na x na matrix
 * (W[ia*nb+ib] = W[ib*nb+ia] after the basis sign of the interleaved
 * convention), one of the odd sector an antisymmetric one.  In either sector
 * sigma needs one same-spin GEMM G = H X, y = hii*x + (G +- G^T), and the
 * mixed term for the upper triangle only, mirrored with the sector's sign --
 * about half the work of the general form.  The mode applies to the
 * thick-restart solves (sbdx_set_restart > 0) and to sbdx_davidson_block
 * (whose columns each lie in one sector; LOWEST keeps both sectors in one
 * basis and returns the lowest roots of H):
 *   SBDX_SYM_OFF     the general form;
 *   SBDX_SYM_EVEN    the even sector: the start vector must be symmetric
 *                    (the Hartree-Fock vector of BasisInitVector is), the
 *                    solve returns the sector's lowest state;
 *   SBDX_SYM_ODD     the odd sector likewise (antisymmetric start vector);
 *   SBDX_SYM_LOWEST  both: the even sector from the start vector's even
 *                    part, then the odd sector from its odd part (or, when
 *                    that is negligible, from (e_ab - e_ba)/sqrt 2 on the
 *                    lowest off-diagonal diagonal element hii[a,b], a < b);
 *                    the odd stage stops once its Ritz value E0 lies above
 *                    the even result by more than 10 |r| (near convergence
 *                    E0 exceeds the eigenvalue it approaches by ~|r|^2/gap),
 *                    and the lower of the two results is returned.
 * The reference's Davidson from the Hartree-Fock vector stays in the even
 * sector up to rounding, so it can only find an odd ground state (found for
 * synthetic instances of 3 and 5 electrons per spin) through rounding noise
 * after many steps; SBDX_SYM_LOWEST returns the ground state of either
 * sector.  Errors with SBDX_ERR_ARG when the string sets differ, and the
 * solve errors when the start vector is outside the chosen sector. */
#define SBDX_SYM_OFF 0
#define SBDX_SYM_EVEN 1
#define SBDX_SYM_ODD 2
#define SBDX_SYM_LOWEST 3
int sbdx_set_symmetry(sbdx_handle_t h, int mode);
/* Per sector (0 even, 1 odd) of the last thick solve: its final energy,
 * sigma products and state (0 not run, 1 converged, 2 not converged, 3 odd
 * stage stopped once above the even result); *chosen the sector returned
 * (1 even, 2 odd, 0 when the solve used the general form). */
int sbdx_sector_stats(sbdx_handle_t h, double *energies, int *products, int *states, int *chosen);
/* sigma of a vector of either spin-flip sector (reference convention) in
 * the symmetric form of that sector, FP64 (fp32 = 0) or FP32 arithmetic; Wb
 * is overwritten.  SBDX_ERR_ARG when the vector lies in neither sector. */
int sbdx_sigma_sym(sbdx_handle_t h, const double *Wk, double *Wb, int fp32);

/* Coarse start of the thick-restart solves (no counterpart in the
 * reference): before the full solve, the start vector restricted to the
 * first m alpha and m beta strings (the strings are sorted, so these hold
 * the lowest excitations) is refined on that m x m sub-problem (same
 * integrals and solver settings, diagonal preconditioner) until its
 * residual is below eps, then lifted back (zero on the other determinants)
 * as the full solve's start.  The full solve still converges to eps on
 * exact FP64 products.  Levels nest: the sub-problem itself starts from its
 * first m/4 strings while m/4 >= 256.  m = 0 (default) switches it off.
 * sbdx_davidson_block uses it too: the lowest nroots roots of the
 * sub-problem (at most 25 block iterations, from the given vectors
 * restricted to it, or from its lowest diagonal elements when they keep
 * less than 1e-3 of their weight there) become the start vectors.
 * Statistics of the last solve's coarse stage: sub-problem products,
 * energy, wall seconds. */
int sbdx_set_coarse_start(sbdx_handle_t h, size_t m, double eps);
int sbdx_coarse_stats(sbdx_handle_t h, int *products, double *energy, double *seconds);

/* hipGraph replay (no counterpart in the reference): with on != 0 the FP32
 * symmetric sigma core (same-spin GEMM, combine, mixed-term kernel on the
 * handle's fixed FP32 buffers) is captured once, after one eager call, and
 * replayed by later symmetric FP32 products (mixed-precision thick solves
 * with sbdx_set_symmetry).  on = 0 drops the graph and runs eagerly. */
int sbdx_set_graphs(sbdx_handle_t h, int on);

/* verbose != 0: the solves print one line per step to stderr, step index,
 * E0 and |r| (as davidson.h:839-851 prints per step); thick-restart solves
 * mark the exact FP64 residual of each restart.  0 (default) prints nothing. */
int sbdx_set_verbose(sbdx_handle_t h, int verbose);

/* Wb = H*Wk evaluated in FP32 (inputs rounded to FP32, result widened):
 * the product the mixed solver applies to correction vectors. */
int sbdx_sigma_fp32(sbdx_handle_t h, const double *Wk, double *Wb);

/*
 * Learned preconditioner (no counterpart in the reference, whose
 * preconditioner is R/(E0-d), davidson.h:861-868): a per-determinant MLP
 * 160 -> 128 -> 128 -> 1 with GELU (tanh form), BF16 weights on MFMA with FP32
 * accumulation, FP32 master weights trained with Adam by the library's own
 * forward/backward kernels.  Features are BF16, sbdx_mlp_nfeat() per row;
 * parameters are FP32, sbdx_mlp_nparam() in the order W1[160][128], b1[128],
 * W2[128][128], b2[128], w3[128], b3 (+3 pad).
 */
int sbdx_mlp_nparam(void);
int sbdx_mlp_nfeat(void);
int sbdx_mlp_set_weights(sbdx_handle_t h, const float *params, size_t n);
int sbdx_mlp_get_weights(sbdx_handle_t h, float *params, size_t n);
/* out[i] = MLP(feat[i]) for n rows (host buffers). */
int sbdx_mlp_forward(sbdx_handle_t h, const uint16_t *feat, size_t n, float *out);
/* Gradient of mean((MLP(feat) - target)^2) over the n rows, one batch. */
int sbdx_mlp_gradient(sbdx_handle_t h, const uint16_t *feat, const float *target, size_t n,
                      float *grad, double *sum_sq_err);
/* Adam (lr, mini-batches of `batch`, shuffled with seed) for `epochs`
 * passes; loss[epoch] receives each epoch's mean squared error. */
int sbdx_mlp_train(sbdx_handle_t h, const uint16_t *feat, const float *target, size_t n, int epochs,
                   int batch, float lr, uint64_t seed, double *loss);
/* Record up to cap (residual, Ritz vector, Ritz value) steps during the next
 * thick-restart solves (cap = 0 stops recording). */
int sbdx_mlp_record(sbdx_handle_t h, int cap);
/* Training pairs from the recorded steps of the last solve, against its
 * converged vector: up to max_pairs rows of features and targets. */
int sbdx_mlp_collect(sbdx_handle_t h, size_t max_pairs, uint64_t seed, uint16_t *feat, float *target,
                     size_t *n_pairs);
/* Preconditioner of thick-restart solves: 0 diagonal, 1 MLP (falls back to
 * the diagonal step for 3 expansions whenever |r| has not dropped by 10%
 * over the last 3 steps).  Counts of each kind in the last solve: */
int sbdx_set_preconditioner(sbdx_handle_t h, int mode);
/* The learned step itself: out[I] = MLP(features of determinant I for the
 * residual R and Ritz value E), R and out indexed I = ia*n_bdet + ib in the
 * handle's alpha-first sign convention (the basis sign is not applied);
 * the solver normalizes R to unit norm before this step. */
int sbdx_mlp_apply(sbdx_handle_t h, const double *R, double E, double *out);
int sbdx_precond_stats(sbdx_handle_t h, long *mlp_steps, long *diag_steps);

/* History of the last solve: trace[2*k] = E0, trace[2*k+1] = |r| of step k,
 * the values the reference prints per Davidson step (davidson.h:839-851). */
int sbdx_trace(sbdx_handle_t h, double *trace, int cap, int *n);

/* Energy <W|H W> of a normalized host vector W (chemistry/tpb/sbdiag.h:431-469). */
int sbdx_energy(sbdx_handle_t h, const double *W, double *energy);

/*
 * 1-/2-RDM of a host vector W, chemistry/tpb/correlation.h:378-662 with the
 * element rules of chemistry/basic/correlation.h:16-229:
 *   one[s*L*L + i + L*j]                           s = 0 alpha, 1 beta
 *   two[(s+2t)*L^4 + i + L*j + L^2*k + L^3*l]
 * one holds 2*L*L doubles and two 4*L^4 doubles (caller-allocated); both are
 * overwritten.
 */
int sbdx_rdm(sbdx_handle_t h, const double *W, double *one, double *two);
/* The RDMs of sbdx_rdm with device pointers, allocating nothing (no
 * counterpart in the reference; for callers that cannot allocate device
 * memory: code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  d_W (n_det doubles, reference convention), d_one (2*L*L) and
 * d_two (4*L^4) are device buffers of the handle's device, which must be
 * current; the outputs are overwritten and the call returns when they are
 * complete.  Working memory comes from a static device array of 128 MiB per
 * device (held from the first call on that device until the process ends)
 * and the handle's own buffers: n_alpha + n_beta + n_alpha*L + L^4 doubles
 * must fit in it with room for 64 pair-table entries (the pair tables are
 * processed in chunks), else SBDX_ERR_ARG names the sizes (norb 58 at
 * 20000 x 20000 strings fits).  Calls on one device share the array: calls
 * that overlap on one device, from any handle or thread, race. */
int sbdx_rdm_device_noalloc(sbdx_handle_t h, const double *d_W, double *d_one, double *d_two);

/*
 * Excitation tables as the reference's TaskHelpers lists
 * (chemistry/tpb/helper.h:185-323): for bra string i, the ket strings of
 * the same spin set that differ by one (kind 1) or two (kind 2) orbitals, in
 * ascending index, with cran = [cr..., an...] spin orbitals (cr bra-only,
 * an ket-only; 2*orb for alpha, 2*orb+1 for beta).  spin 0 alpha, 1 beta.
 * sbdx_table_count gives the entry count; export fills offsets[n+1],
 * kets[count] and cran[count*2*kind].
 */
int sbdx_table_count(sbdx_handle_t h, int spin, int kind, size_t *count);
int sbdx_table_export(sbdx_handle_t h, int spin, int kind, uint64_t *offsets,
                      uint32_t *kets, int32_t *cran);

/* Per-kernel timing with HIP events on the handle's stream.
 * enable != 0 starts recording (and clears earlier records). */
int sbdx_profile(sbdx_handle_t h, int enable);
/* Number of distinct kernels recorded so far. */
int sbdx_profile_count(sbdx_handle_t h, int *n);
/* Name, launch count and summed duration (ms) of recorded kernel idx. */
int sbdx_profile_get(sbdx_handle_t h, int idx, char *name, size_t name_cap,
                     long *launches, double *total_ms);

/* Algorithmic work of one sigma for this instance (roofline numerators):
 * counts[0..3] = alpha singles, alpha doubles, beta singles, beta doubles
 * table entries; counts[4] = sum over alpha rows of (alpha singles x sum of
 * beta singles) = the number of mixed terms; counts[5] = n_det. */
int sbdx_work_counts(sbdx_handle_t h, double *counts);

#ifdef __cplusplus
}
#endif
#endif /* SBDX_H */
