/*
 * powerspectra_hip.h -- C ABI of the MI355X (gfx950) mode-coupling and
 * covariance engine that replaces the inner loops of PowerSpectra.jl.
 *
 * Every entry point here is what a Julia `ccall` (or Python ctypes) binds:
 * plain pointers, sizes and an int status.  No exceptions, no C++ or torch
 * types cross the boundary.
 *
 * Conventions shared by all entry points
 *   - FP64 everywhere on the boundary.
 *   - Matrices are column-major.  Element (l1, l2) of an lmin:lmax block is at
 *     M[(l1 - lmin) + (l2 - lmin) * ldM]  (parent(SpectralArray) layout,
 *     reference src/spectralarray.jl:5-45, src/modecoupling.jl:90-91).
 *   - Spectra (W, signal, noise ratios) are 0-based: W[l] for l = 0 .. LW-1
 *     (SpectralVector, reference src/spectralarray.jl:17).
 *   - Host-pointer entry points copy in and out; no pointer is retained.
 *     They keep their device copies in a per-context transfer buffer
 *     between calls, and the transforms their work arrays in a per-context
 *     work buffer (both counted in psx_ctx_bytes_in_use, given back when a
 *     later allocation would otherwise exceed the budget).  Mode-coupling
 *     and covariance outputs over 32 MB are computed in l1 bands, the
 *     columns each band completes copied to the host while the later bands
 *     compute.
 *     `_d` entry points take device pointers and enqueue on the context's
 *     stream; they return after the work is enqueued (call
 *     psx_ctx_synchronize to wait).
 *   - Calls on one context are serialised by the context.  One context owns
 *     one HIP stream on one device.
 */
#ifndef POWERSPECTRA_HIP_H
#define POWERSPECTRA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ----------------------------------------------------- */
enum psx_status {
  PSX_OK = 0,
  PSX_ERR_ARGUMENT = 1,   /* bad spec / kind / size  (Julia ArgumentError)       */
  PSX_ERR_AXES = 2,       /* axis mismatch           (Julia AssertionError)      */
  PSX_ERR_SINGULAR = 3,   /* zero pivot in LU        (LinearAlgebra.SingularException) */
  PSX_ERR_DOMAIN = 4,     /* sqrt of a negative spectrum product (DomainError)   */
  PSX_ERR_BUDGET = 5,     /* request exceeds the context's HBM budget            */
  PSX_ERR_DEVICE = 6,     /* HIP / rocSOLVER failure                             */
  PSX_ERR_NOT_READY = 7   /* serving: ticket not finished yet                    */
};

/* ---- mode-coupling kinds (reference src/modecoupling.jl:78-159) -------- */
enum psx_mcm_kind {
  PSX_MCM_00 = 0,  /* inner_mcm⁰⁰!  (TT)                     :78-95   */
  PSX_MCM_02 = 1,  /* inner_mcm⁰²!  (TE, ET, TB, BT)         :99-119  */
  PSX_MCM_PP = 2,  /* inner_mcm⁺⁺!  (spin-2, even l1+l2+l3)  :123-139 */
  PSX_MCM_MM = 3   /* inner_mcm⁻⁻!  (spin-2, odd  l1+l2+l3)  :143-159 */
};

/* ---- covariance kinds (reference src/covariance.jl:92-446) ------------- */
enum psx_cov_kind {
  PSX_COV_TTTT = 0,        /* loop_covTTTT!         :92-122  K=8 */
  PSX_COV_EEEE = 1,        /* loop_covEEEE!         :153-183 K=8 */
  PSX_COV_TTTE = 2,        /* loop_covTTTE!         :208-235 K=4 */
  PSX_COV_TETE = 3,        /* loop_covTETE!         :261-302 K=5 */
  PSX_COV_TEEE = 4,        /* loop_covTEEE!         :337-372 K=4 (planck=false) */
  PSX_COV_TEEE_PLANCK = 5, /* loop_covTEEE_planck!  :376-402 K=4 (default)      */
  PSX_COV_TTEE = 6         /* loop_covTTEE!         :422-446 K=2 */
};

/* Number of W spectra, signal spectra and noise-ratio spectra each covariance
 * kind takes, in the argument order of the reference loop (see
 * psx_cov).  Returns PSX_ERR_ARGUMENT for an unknown kind. */
int psx_cov_arity(int kind, int* nW, int* nspectra, int* nratios);

enum psx_precision {
  PSX_FP64 = 0,      /* FP64 on the vector ALUs (the reference's arithmetic)     */
  PSX_MIXED = 1,     /* the contraction against the W spectra on the int8 matrix
                        cores: the 3j products split exactly into six int8
                        digits, the W spectra into seven, 27 digit products
                        accumulated exactly in int32 and combined in FP64
                        (family products and epilogues stay FP64); error near
                        1e-14 of the element's terms                          */
  PSX_FP64_MFMA = 2  /* the same contraction on the FP64 matrix cores
                        (v_mfma_f64_16x16x4): FP64 products and sums            */
};

/* ---- context ----------------------------------------------------------- */
typedef struct psx_ctx psx_ctx;

/* Create a context on HIP device `device`.  `hbm_budget` bounds the bytes the
 * context may hold for workspaces and resident spectra (0 = 90% of free HBM). */
int psx_ctx_create(int device, size_t hbm_budget, psx_ctx** out);
/* Waits for the context's work and frees everything it holds.  Refused with
 * PSX_ERR_ARGUMENT (nothing destroyed) while a server created on this context
 * is alive: destroy servers first (psx_server_destroy). */
int psx_ctx_destroy(psx_ctx* ctx);
int psx_ctx_synchronize(psx_ctx* ctx);
/* Human-readable message for the last error on this context (never NULL). */
const char* psx_ctx_last_error(psx_ctx* ctx);
/* Bytes of device memory the context holds right now. */
size_t psx_ctx_bytes_in_use(psx_ctx* ctx);

/* Kernel timing: when enabled, every kernel launch is bracketed by HIP events
 * on the context stream and its duration accumulated under the kernel's name. */
int psx_ctx_set_timing(psx_ctx* ctx, int enable);
/* Fetch accumulated time (ms) and launch count for kernel `name`; resets it
 * when `reset` != 0. Unknown names report zero launches. */
int psx_ctx_kernel_time(psx_ctx* ctx, const char* name, int reset,
                        double* total_ms, int64_t* launches);

/* Device memory owned by the context (for resident inputs / outputs). */
int psx_dev_alloc(psx_ctx* ctx, size_t bytes, void** dptr);
int psx_dev_free(psx_ctx* ctx, void* dptr);
int psx_memcpy_h2d(psx_ctx* ctx, void* dst, const void* src, size_t bytes);
int psx_memcpy_d2h(psx_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ---- mode-coupling matrices ------------------------------------------- */
/* One block.  Replaces inner_mcm⁰⁰!/⁰²!/⁺⁺!/⁻⁻!(M, V) (reference
 * src/modecoupling.jl:78-159), i.e. fills M[l1,l2] = (2l2+1) Ξ and
 * M[l2,l1] = (2l1+1) Ξ for lmin <= l1 <= l2 <= lmax, with Ξ the projector of
 * src/modecoupling.jl:3-66 summed over l3 <= min(l1+l2, LW-1).
 * The reference passes V = alm2cl(...)[0:lmax], i.e. LW = lmax+1
 * (src/modecoupling.jl:197); LW is a free parameter here.
 * Every mode-coupling and covariance entry point accepts lmax <= 8191: up
 * to there the integer polynomials of the closed-form 3j products are exact
 * doubles, so each term carries a few ulp (PSX_ERR_ARGUMENT above). */
int psx_mcm(psx_ctx* ctx, int kind, int lmin, int lmax,
            const double* W, int64_t LW, double* M, int64_t ldM);
int psx_mcm_d(psx_ctx* ctx, int kind, int lmin, int lmax,
              const double* dW, int64_t LW, double* dM, int64_t ldM);
/* psx_mcm_d without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  Same kernels, launch geometry and
 * bits as psx_mcm_d; everything beyond dW and dM lives in one static device
 * array of 4 MiB per device, set up by psx_ctx_create and held until the
 * process ends.  Limits: lmax <= 8191 (PSX_ERR_ARGUMENT above).  Every call
 * on a device shares that device's array: calls that overlap on one device
 * (several contexts or threads) race.  Enqueued on the context stream. */
int psx_mcm_d_noalloc(psx_ctx* ctx, int kind, int lmin, int lmax,
                      const double* dW, int64_t LW, double* dM, int64_t ldM);

/* Binned mode-coupling matrix P M Q (the binning of reference src/util.jl:3-18
 * applied to one block of psx_mcm): P is nbl x n (leading dimension nbl),
 * Q is n x nbr (leading dimension n), n = lmax-lmin+1, both over l = lmin..lmax;
 * out is nbl x nbr (leading dimension ldo).  M stays on the device: only the
 * binned matrix is copied back. */
int psx_mcm_binned(psx_ctx* ctx, int kind, int lmin, int lmax, const double* W, int64_t LW,
                   const double* P, int nbl, const double* Q, int nbr, double* out,
                   int64_t ldo);

/* Fused spin-2 pair: one pass fills M⁺⁺ and M⁻⁻ from the same W (the two
 * reference passes inner_mcm⁺⁺! and inner_mcm⁻⁻! of src/modecoupling.jl:216-217
 * or :238-239 share their (0 -2 2) family). */
int psx_mcm_pm(psx_ctx* ctx, int lmin, int lmax, const double* W, int64_t LW,
               double* Mpp, double* Mmm, int64_t ldM);
int psx_mcm_pm_d(psx_ctx* ctx, int lmin, int lmax, const double* dW, int64_t LW,
                 double* dMpp, double* dMmm, int64_t ldM);

/* Full polarised set in one pass over (l1, l2): TT from W_TT, TE from W_TE,
 * ET from W_ET, ++ and -- from W_PP -- the five mcm() calls made by
 * maskedalm2spectra (reference src/modecoupling.jl:348-377).  Any W pointer
 * may be NULL to skip that block (its M pointer is then ignored). */
int psx_mcm_set(psx_ctx* ctx, int lmin, int lmax, int64_t LW,
                const double* W_TT, const double* W_TE, const double* W_ET,
                const double* W_PP,
                double* M_TT, double* M_TE, double* M_ET, double* M_PP,
                double* M_MM, int64_t ldM);
int psx_mcm_set_d(psx_ctx* ctx, int lmin, int lmax, int64_t LW,
                  const double* dW_TT, const double* dW_TE, const double* dW_ET,
                  const double* dW_PP,
                  double* dM_TT, double* dM_TE, double* dM_ET, double* dM_PP,
                  double* dM_MM, int64_t ldM);

/* ---- coupled covariance ------------------------------------------------ */
/* Replaces loop_cov*!(C, spectra..., ratios..., W1..WK) (reference
 * src/covariance.jl:92-446).  `spectra`, `ratios` and `W` are arrays of
 * pointers in exactly the reference loop's argument order; psx_cov_arity
 * gives their counts.  Spectra and ratios are 0-based with length >= lmax+1,
 * W has length LW (reference: workspace.lmax+1).  Fills C[l1,l2] = C[l2,l1]
 * for lmin <= l1 <= l2 <= lmax.  `precision`: PSX_FP64, or PSX_MIXED /
 * PSX_FP64_MFMA through the matrix-core pass of psx_cov_batch (not TETE). */
int psx_cov(psx_ctx* ctx, int kind, int lmin, int lmax,
            const double* const* spectra, const double* const* ratios,
            const double* const* W, int64_t LW, int precision,
            double* C, int64_t ldC);
int psx_cov_d(psx_ctx* ctx, int kind, int lmin, int lmax,
              const double* const* dspectra, const double* const* dratios,
              const double* const* dW, int64_t LW, int precision,
              double* dC, int64_t ldC);

/* ---- batches: many W spectra against one family in one pass ------------ */
/* K blocks of one mode-coupling kind from K W spectra (K = 1..64), the
 * result of K psx_mcm calls (reference inner_mcm*! of
 * src/modecoupling.jl:78-159 once per W): every family product is computed
 * once and contracted against all K spectra.  precision: PSX_FP64 (vector
 * FP64, groups of up to 8 spectra per launch), PSX_MIXED (int8 matrix cores,
 * exact digit products; the tests bound the error per element at 1e-9) or
 * PSX_FP64_MFMA (FP64 matrix cores).  W[k] has LW
 * entries, M[k] is n x n with leading dimension ldM. */
int psx_mcm_batch(psx_ctx* ctx, int kind, int lmin, int lmax, int K, const double* const* W,
                  int64_t LW, int precision, double* const* M, int64_t ldM);
int psx_mcm_batch_d(psx_ctx* ctx, int kind, int lmin, int lmax, int K,
                    const double* const* dW, int64_t LW, int precision, double* const* dM,
                    int64_t ldM);
/* nblk covariance blocks of one kind (the result of nblk psx_cov calls,
 * reference loop_cov*! of src/covariance.jl:92-446): spectra, ratios and W
 * hold the blocks' arguments one block after another, each block in the
 * reference loop's order with psx_cov_arity's counts (nblk * nS spectra,
 * nblk * nR ratios, nblk * nW W spectra); C[b] is block b.  The matrix-core
 * precisions contract up to 64 W spectra per pass (8 TTTT or EEEE blocks,
 * 16 TTTE, TEEE or TEEE-Planck blocks, 16 TTEE blocks) and take every kind
 * but TETE, whose two families need two passes (PSX_ERR_ARGUMENT). */
int psx_cov_batch(psx_ctx* ctx, int kind, int lmin, int lmax, int nblk,
                  const double* const* spectra, const double* const* ratios,
                  const double* const* W, int64_t LW, int precision, double* const* C,
                  int64_t ldC);
int psx_cov_batch_d(psx_ctx* ctx, int kind, int lmin, int lmax, int nblk,
                    const double* const* dspectra, const double* const* dratios,
                    const double* const* dW, int64_t LW, int precision, double* const* dC,
                    int64_t ldC);

/* ---- FP64 dense linear algebra ----------------------------------------- */
/* decouple_covmat (reference src/covariance.jl:8-14): out = B1^{-1} Y B2^{-T},
 * all n x n column-major with leading dimension n.  Y is not modified.  When
 * B1 and B2 are the same pointer the matrix is factored once. */
int psx_decouple_covmat(psx_ctx* ctx, int n, const double* Y, const double* B1,
                        const double* B2, double* out);
/* `A \ B` (reference src/blockspectralmatrix.jl:101-129): LU with partial
 * pivoting of the n x n matrix A, then solve for nrhs right-hand sides in B
 * (n x nrhs, overwritten with the solution).  A is not modified. */
int psx_lu_solve(psx_ctx* ctx, int n, const double* A, int64_t lda,
                 double* B, int64_t ldb, int nrhs);
/* The same on device buffers: dA is overwritten by its LU factors and dB by
 * the solution; enqueued on the context stream, returns when done. */
int psx_lu_solve_d(psx_ctx* ctx, int n, double* dA, int64_t lda, double* dB, int64_t ldb,
                   int nrhs);
/* The joint EE/BB (sign = +1) or EB/BE (sign = -1) decoupling solve
 * [M++ s M--; s M-- M++] X = B of maskedalm2spectra (reference
 * src/modecoupling.jl:362-377, the blocks of mcm(:EE_BB) and mcm(:EB_BE),
 * :210-224,235-244): Mpp, Mmm are n x n (leading dimension ldM), B is 2n x nrhs
 * (leading dimension ldb >= 2n) and is overwritten by X.  The 2n x 2n matrix
 * is assembled on the device, so only M++ and M-- cross the bus; sign = -1
 * solves the sign = +1 matrix against D B and returns D X, D = diag(I, -I),
 * which is exact.  LU with partial pivoting as psx_lu_solve. */
int psx_lu_solve_pm(psx_ctx* ctx, int n, int sign, const double* Mpp, const double* Mmm,
                    int64_t ldM, double* B, int64_t ldb, int nrhs);
/* The decoupling half of maskedalm2spectra (reference
 * src/modecoupling.jl:348-377): the five blocks TT, TE (= TB), ET (= BT),
 * ++ and -- from the mask cross-spectra W_TT, W_TE, W_ET, W_PP (length LW,
 * as psx_mcm_set), then the nine spectra M \ pCl.  pcl holds the nine
 * pseudo-spectra for l = lmin..lmax, n = lmax - lmin + 1 each, consecutive
 * in the order TT, TE, TB, ET, BT, EE, BB, EB, BE; out receives the
 * decoupled spectra in the same layout.  The blocks never leave the device:
 * TT, TE (for TE and TB), ET (for ET and BT) and the 2n x 2n
 * [M++ M--; M-- M++] (for EE/BB and, through D = diag(I, -I), EB/BE) are
 * factored at the same time on separate streams.  LU with partial pivoting
 * as psx_lu_solve; a zero pivot returns PSX_ERR_SINGULAR. */
int psx_master_decouple(psx_ctx* ctx, int lmin, int lmax, int64_t LW, const double* W_TT,
                        const double* W_TE, const double* W_ET, const double* W_PP,
                        const double* pcl, double* out);
/* `lu(A)` (reference src/blockspectralmatrix.jl:89-99, Julia's LAPACK getrf
 * with partial pivoting): LU (n x n, leading dimension ldlu) receives the
 * unit-lower L below the diagonal and U on and above it, ipiv[n] the 1-based
 * row interchanges (row j was swapped with row ipiv[j] at step j).  A zero
 * pivot returns PSX_ERR_SINGULAR (Julia SingularException) after writing both.
 * Ties in |pivot| go to the lower row index.  The factorisation behind
 * psx_lu_solve and psx_decouple_covmat. */
int psx_lu_factor(psx_ctx* ctx, int n, const double* A, int64_t lda, double* LU,
                  int64_t ldlu, int* ipiv);

/* ---- spherical-harmonic transforms (mask -> a_lm -> W_l front end) ----- */
/* Analysis of `nmaps` ring-ordered HEALPix maps (12 nside^2 doubles each,
 * consecutive) into a_lm, 0 <= m <= l <= lmax, m-major (index
 * m (2 lmax + 1 - m) / 2 + l), complex as interleaved (re, im) doubles,
 * (lmax+1)(lmax+2)/2 coefficients per map, with `niter` Jacobi refinements.
 * Replaces Healpix.jl's map2alm as called by mcm(spec, mask1, mask2) and the
 * covariance workspace (reference src/modecoupling.jl:250-251,
 * src/workspace.jl:155,163).  nside a power of two <= 2048,
 * lmax <= min(4 nside, 8191). */
int psx_map2alm(psx_ctx* ctx, int nside, int lmax, int niter, int nmaps,
                const double* maps, double* alms);
/* Synthesis (Healpix.jl alm2map): the inverse of one analysis step. */
int psx_alm2map(psx_ctx* ctx, int nside, int lmax, int nmaps, const double* alms,
                double* maps);
/* The mask -> W front end in one call: map2alm (niter refinements) of nmaps
 * ring-ordered maps (1..8, consecutive) and all their cross-spectra
 * cls[(i nmaps + j)(lmax+1) + l] = alm2cl(a_i, a_j)[l]
 *   = (2l+1)^-1 [Re(a_l0 b_l0*) + 2 sum_{m>=1} Re(a_lm b_lm*)],
 * i.e. `alm2cl(map2alm(m_i), map2alm(m_j))` of mcm(spec, mask1, mask2) and the
 * covariance workspace (reference src/modecoupling.jl:197,251,255,
 * src/workspace.jl:202), with the alms kept on the device.  Limits as
 * psx_map2alm. */
int psx_mask_spectra(psx_ctx* ctx, int nside, int lmax, int niter, int nmaps,
                     const double* maps, double* cls);
/* Healpix.jl alm2cl of given coefficient sets, as maskedalm2spectra applies it
 * to the masked maps' and the masks' alms (reference
 * src/modecoupling.jl:355,363-364): alms[i] points at nalm complex a_lm of set i
 * (1..8 sets, layout as psx_map2alm, each set anywhere in host memory), and
 * cls[(i nsets + j)(lmax+1) + l] = alm2cl(a_i, a_j)[l] for every i, j, summed
 * as in psx_mask_spectra.  lmax 0..8191. */
int psx_alm2cl(psx_ctx* ctx, int lmax, int nsets, const double* const* alms, double* cls);
/* The spectra `master` decouples (reference src/modecoupling.jl:320-330: the
 * masking and map2alm calls; the alm2cl of mcm's masks, :197, and of the
 * pseudo-spectra in maskedalm2spectra, :355,363-364), from two polarised maps
 * and their masks in one call: maps[0..5] = I1, Q1, U1, I2, Q2, U2 and
 * masks[0..3] = T1, P1, T2, P2, ring order, 12 nside^2 doubles each, anywhere
 * in host memory (a mask address given twice is copied and transformed once).
 * On the device: I times its T mask and Q, U times the P mask (skipped when
 * already_masked != 0), the polarised analysis of both maps and the scalar
 * analysis of the masks (niter refinements each, as psx_map2alm_pol and
 * psx_map2alm), and their cross-spectra, summed as in psx_mask_spectra:
 *   pcl[(x 3 + y)(lmax+1) + l] = alm2cl(X of map 1, Y of map 2)[l], x, y over T, E, B;
 *   W[(a 2 + b)(lmax+1) + l]   = alm2cl(mask a of map 1, mask b of map 2)[l],
 *                                a, b over T, P.
 * Only the spectra come back to the host.  Limits as psx_map2alm. */
int psx_master_spectra(psx_ctx* ctx, int nside, int lmax, int niter, const double* const* maps,
                       const double* const* masks, int already_masked, double* pcl, double* W);
/* Polarised analysis (Healpix.jl map2alm of a PolarizedHealpixMap, as master
 * calls it, reference src/modecoupling.jl:328-329): per map the I, Q, U rings
 * (3 x 12 nside^2 doubles, consecutive) into T, E, B a_lm (3 x nalm complex,
 * layout as psx_map2alm), HEALPix convention Q +- iU = sum _{+-2}a _{+-2}Y,
 * a_E = -(_2a + _-2a)/2, a_B = i(_2a - _-2a)/2; `niter` refinements of both
 * parts.  a_E, a_B are zero for l < 2.  Limits as psx_map2alm. */
int psx_map2alm_pol(psx_ctx* ctx, int nside, int lmax, int niter, int nmaps,
                    const double* iqu, double* teb);
/* Polarised synthesis: T, E, B a_lm to I, Q, U maps. */
int psx_alm2map_pol(psx_ctx* ctx, int nside, int lmax, int nmaps, const double* teb,
                    double* iqu);

/* ---- QuickPol beam matrix --------------------------------------------- */
/* quickpolΞ! (reference src/beam.jl:72-101, with Ξsum :17-28 and
 * specrowrange :57-61) on a BandedMatrix in BandedMatrices.jl band storage:
 * (bl + bu + 1) x (lmax + 1), column-major, element (row l'', column l) at
 * band[(bu + l'' - l) + l (bl + bu + 1)].  For 2 <= l'' <= lmax and
 * max(2, l'' - bl) <= l <= min(lmax, l'' + bu):
 *   band(l'', l) = (-1)^(s1+s2+nu1+nu2) sum_{l'} W[l'] f1(l') f2(l'),
 *   f_k = WignerF(l, l'', -s_k, -nu_k) = (l' l l''; s_k+nu_k, -s_k, -nu_k);
 * every other stored entry (rows or columns 0, 1) is multiplied by the same
 * sign, as the reference's `𝚵 .*= sgn` does, so `band` is read and written.
 * W is 0-based with LW entries (quickpolW of the two scan-weight alms);
 * terms with l' >= LW count as zero (the reference reads W[l'] unchecked).
 * A family with |s| > l or |nu| > l'' is zero.  lmax <= 65535, bandwidths
 * 0..lmax, spins within -8..8. */
int psx_quickpol_xi(psx_ctx* ctx, int nu1, int nu2, int s1, int s2, int lmax, int bl, int bu,
                    const double* W, int64_t LW, double* band);
/* The same on device buffers, enqueued on the context stream (returns when
 * done). */
int psx_quickpol_xi_d(psx_ctx* ctx, int nu1, int nu2, int s1, int s2, int lmax, int bl,
                      int bu, const double* dW, int64_t LW, double* dband);

/* ---- serving: a persistent request loop ------------------------------- */
/* The reference has no serving layer (its callers run mcm / coupledcov one
 * call at a time); these entry points batch those same calls for callers that
 * issue many of them.  A server owns `slots` worker threads, each with its own
 * HIP stream, and a device arena of `arena_bytes` from which every request's
 * outputs and scratch are taken.  Requests are admitted in arrival order as
 * soon as a slot is free and the arena can hold the request; admitted requests
 * run concurrently on the device.  Spectra (mask W, signal spectra, noise
 * ratios) are made resident once under a caller-chosen 64-bit key and are
 * referenced by key.  Results are bit-identical to the single-call entry points
 * (same kernels, same summation order), alone or in a mix. */
typedef struct psx_server psx_server;

enum psx_request_type {
  PSX_REQ_MCM = 0,     /* one block, kind = psx_mcm_kind, W[0]        (psx_mcm)     */
  PSX_REQ_MCM_PM = 1,  /* ++ and -- from W[0]; out[0] = ++, out[1] = -- (psx_mcm_pm) */
  PSX_REQ_MCM_SET = 2, /* TT, TE, ET, ++, -- from W[0..3] = TT, TE, ET, PP (psx_mcm_set) */
  PSX_REQ_COV = 3      /* one covariance block, kind = psx_cov_kind     (psx_cov)     */
};

typedef struct psx_request {
  int type;            /* psx_request_type                                         */
  int kind;            /* psx_mcm_kind (PSX_REQ_MCM) or psx_cov_kind (PSX_REQ_COV)  */
  int lmin, lmax;
  int64_t LW;          /* W length used; every W key must hold at least LW values  */
  uint64_t W[8];       /* resident keys, count as for the matching entry point     */
  uint64_t spectra[4]; /* PSX_REQ_COV: resident keys, length >= lmax + 1           */
  uint64_t ratios[4];
  double* out[5];      /* host outputs (ld >= lmax-lmin+1), or all NULL to keep the */
  int64_t ld;          /* result on the device until psx_fetch / psx_release        */
} psx_request;

/* slots: concurrent requests (1..16); arena_bytes: device memory for requests
 * (taken from the context budget); lmax_max: largest lmax accepted (<= 8191).
 * The server lives on its context: psx_ctx_destroy refuses while it exists. */
int psx_server_create(psx_ctx* ctx, int slots, size_t arena_bytes, int lmax_max,
                      psx_server** out);
/* Waits for admitted work, drops queued requests, frees everything. */
int psx_server_destroy(psx_server* s);
/* Admission order.  PSX_SCHED_FIFO (default): arrival order.
 * PSX_SCHED_DEADLINE: earliest deadline first, deadline = arrival time +
 * slack x the request's estimated run time alone (from its size and kind), so
 * small requests are not held behind large ones while no request waits
 * indefinitely.  Either way the request chosen waits for arena space before
 * any other is admitted. */
enum psx_sched { PSX_SCHED_FIFO = 0, PSX_SCHED_DEADLINE = 1 };
int psx_server_set_policy(psx_server* s, int policy, double slack);
/* Copy `len` doubles to the device under `key` (replacing a previous entry of
 * that key that no pending request uses). */
int psx_resident_put(psx_server* s, uint64_t key, const double* values, int64_t len);
/* PSX_ERR_ARGUMENT while a queued or running request references the key. */
int psx_resident_evict(psx_server* s, uint64_t key);
/* Validate and enqueue; *ticket identifies the request.  PSX_ERR_BUDGET if the
 * request can never fit in the arena. */
int psx_submit(psx_server* s, const psx_request* rq, uint64_t* ticket);
/* *done = 1 once the request has finished (successfully or not). */
int psx_poll(psx_server* s, uint64_t ticket, int* done);
/* Block until the request has finished; *status receives its own status,
 * times (if not NULL) {seconds queued, seconds running incl. copy-out, total}. */
int psx_wait(psx_server* s, uint64_t ticket, int* status, double* times);
/* Copy a device-held result (out[] all NULL at submit) into host arrays. */
int psx_fetch(psx_server* s, uint64_t ticket, double* const* out, int64_t ld);
/* Forget a finished ticket and free its device result, if any. */
int psx_release(psx_server* s, uint64_t ticket);
/* Message of the last failed server call or request status (thread-local). */
const char* psx_server_last_error(psx_server* s);

#ifdef __cplusplus
}
#endif
#endif /* POWERSPECTRA_HIP_H */
