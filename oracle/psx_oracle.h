/*
 * psx_oracle.h -- CPU oracle for PowerSpectra.jl's mode-coupling and
 * covariance hot path.
 *
 * TEST INFRASTRUCTURE.  This library is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  It restates the reference algorithm (recurrence-generated Wigner
 * families, the Ξ projector sums, the threaded inner loops, LU decoupling) in
 * plain C, threaded with OpenMP `schedule(dynamic)` over l1 rows as the
 * reference does with ThreadPools.@qthreads.
 *
 * How each function is pinned (checks live in tests/):
 *   psxo_wigner_family     sympy exact 3j values (tests/golden), sum rule,
 *                          long-double self-consistency, sign at nmax.
 *   psxo_mcm               brute force from the sympy fixtures; full-sky and
 *                          flat-W closed forms; symmetry; the reference's own
 *                          NaMaster diagonal fixtures (W fitted from the TT
 *                          diagonal predicts the EE and TE diagonals).
 *   psxo_cov               full-sky Knox closed form; symmetry; brute force
 *                          from sympy fixtures at small lmax; TEEE Planck
 *                          approximation within 1% of exact at l >= 29.
 *   psxo_decouple_covmat   the reference KAT (test/test_covmat.jl:11-17) and
 *                          the C = M C_dec M^T invariant (:115-127);
 *   psxo_lu_solve          numpy.linalg.solve.
 *   psxo_quickpol_xi       general-spin families against the sympy fixtures
 *                          (spins up to 4); for s = nu = 0 it equals the
 *                          pinned TT mode-coupling sums (W -> 4 pi W/(2l+1));
 *                          brute force from sympy at small lmax; band layout
 *                          and the (-1)^(s1+s2+nu1+nu2) sign by hand.
 * Spin-2 rows/columns with l < 2 are unpinned (the reference's value comes
 * from WignerFamilies.jl on invalid |m| > l); the oracle returns zeros there.
 */
#ifndef PSX_ORACLE_H
#define PSX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* f(j) = ( j l1 l2 ; -m1-m2 m1 m2 ) for j = nmin..l1+l2 into out[0..];
 * returns the count (or -1 if cap is too small). */
int psxo_wigner_family(int l1, int l2, int m1, int m2, double* out, int cap,
                       int* nmin);
/* Same in long double; each value returned as hi + lo doubles. */
int psxo_wigner_family_ld(int l1, int l2, int m1, int m2, double* hi,
                          double* lo, int cap, int* nmin);

/* inner_mcm kinds: psx_mcm_kind of include/powerspectra_hip.h.
 * M is column-major (lmax-lmin+1)^2 with leading dimension ldM.
 * precision: 0 = double, 1 = long double (oracle-of-oracle). */
int psxo_mcm(int kind, int lmin, int lmax, const double* W, int64_t LW,
             double* M, int64_t ldM, int nthreads, int precision);

/* Only rows l1 in [l1_lo, l1_hi] of the upper triangle (l2 >= l1), written
 * to both (l1,l2) and (l2,l1); other entries of M are left untouched. */
int psxo_mcm_rows(int kind, int lmin, int lmax, int l1_lo, int l1_hi,
                  const double* W, int64_t LW, double* M, int64_t ldM,
                  int nthreads, int precision);

/* loop_cov* kinds: psx_cov_kind. Argument order as psx_cov. */
int psxo_cov(int kind, int lmin, int lmax, const double* const* spectra,
             const double* const* ratios, const double* const* W, int64_t LW,
             double* C, int64_t ldC, int nthreads);

int psxo_cov_rows(int kind, int lmin, int lmax, int l1_lo, int l1_hi,
                  const double* const* spectra, const double* const* ratios,
                  const double* const* W, int64_t LW, double* C, int64_t ldC,
                  int nthreads);

/* QuickPol Ξ (reference src/beam.jl:72-101) into band storage
 * band[(bu + i - j) + j (bl + bu + 1)], i = l'' (row), j = l (column);
 * in/out: entries outside rows/columns >= 2 are only multiplied by the sign. */
int psxo_quickpol_xi(int nu1, int nu2, int s1, int s2, int lmax, int bl, int bu,
                     const double* W, int64_t LW, double* band, int nthreads);

/* out = B1^{-1} Y B2^{-T} (column-major n x n). */
int psxo_decouple_covmat(int n, const double* Y, const double* B1,
                         const double* B2, double* out);
/* Solve A X = B (A n x n, B n x nrhs, column-major, overwritten). */
int psxo_lu_solve(int n, const double* A, double* B, int nrhs);

#ifdef __cplusplus
}
#endif
#endif
