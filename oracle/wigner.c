/*
 * CPU oracle: Wigner families, Ξ projectors, inner MCM loops, covariance
 * loops and LU decoupling.  TEST INFRASTRUCTURE -- see psx_oracle.h.
 */
#include "psx_oracle.h"
#include "../include/powerspectra_hip.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define SG_A(s) CAT(sg_A, s)
#define SG_Y(s) CAT(sg_Y, s)
#define wigner_family(s) CAT(wigner_family, s)

#define REAL double
#define SFX _d
#define RSQRT sqrt
#define RABS fabs
#include "wigner_impl.h"
#undef REAL
#undef SFX
#undef RSQRT
#undef RABS

#define REAL long double
#define SFX _ld
#define RSQRT sqrtl
#define RABS fabsl
#include "wigner_impl.h"
#undef REAL
#undef SFX
#undef RSQRT
#undef RABS

int psxo_wigner_family(int l1, int l2, int m1, int m2, double* out, int cap,
                       int* nmin) {
  if (l1 < 0 || l2 < 0) return -1;
  if (l1 + l2 + 1 > cap) return -1;
  return wigner_family_d(l1, l2, m1, m2, out, nmin);
}

int psxo_wigner_family_ld(int l1, int l2, int m1, int m2, double* hi,
                          double* lo, int cap, int* nmin) {
  if (l1 < 0 || l2 < 0) return -1;
  if (l1 + l2 + 1 > cap) return -1;
  long double* tmp = (long double*)malloc(sizeof(long double) * (l1 + l2 + 1));
  int n = wigner_family_ld(l1, l2, m1, m2, tmp, nmin);
  for (int k = 0; k < n; ++k) {
    hi[k] = (double)tmp[k];
    lo[k] = (double)(tmp[k] - (long double)hi[k]);
  }
  free(tmp);
  return n;
}

/* ------------------------------------------------------------------------
 * Ξ projectors, reference src/modecoupling.jl:3-66.  `w` holds the family
 * product over l3 = nmin..nmax (w[l3 - nmin]); W is 0-based with length LW.
 * parity: 0 = all l3 (Ξ_TT), 1 = l1+l2+l3 even (Ξ_EE, Ξ_TE), 2 = odd (Ξ_EB).
 * ---------------------------------------------------------------------- */
#define DEFINE_XI(NAME, T)                                                   \
  static T NAME(const double* W, int64_t LW, const T* w, int nmin, int nmax, \
                int l1, int l2, int parity) {                                \
    T xi = 0;                                                                \
    int start = nmin > 0 ? nmin : 0;                                         \
    int64_t end = nmax < LW - 1 ? nmax : LW - 1;                             \
    int step = 1;                                                            \
    if (parity == 1) {                                                       \
      if ((l1 + l2 + start) & 1) start += 1;                                 \
      step = 2;                                                              \
    } else if (parity == 2) {                                                \
      if (((l1 + l2 + start) & 1) == 0) start += 1;                          \
      step = 2;                                                              \
    }                                                                        \
    for (int64_t l3 = start; l3 <= end; l3 += step)                          \
      xi += (T)(2 * l3 + 1) * w[l3 - nmin] * (T)W[l3];                       \
    return xi / (T)(4.0L * 3.14159265358979323846264338327950288L);         \
  }
DEFINE_XI(xi_d, double)
DEFINE_XI(xi_ld, long double)

/* Julia's sqrt throws DomainError on a negative argument; record it. */
static double sq_checked(double x, int* err) {
  if (x < 0) {
    *err = 1;
    return NAN;
  }
  return sqrt(x);
}

static int nthreads_or_default(int n) {
#ifdef _OPENMP
  return n > 0 ? n : omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

/* inner_mcm⁰⁰!/⁰²!/⁺⁺!/⁻⁻!, reference src/modecoupling.jl:78-159 */
#define DEFINE_MCM(NAME, T, FAMILY, XI)                                       \
  static int NAME(int kind, int lmin, int lmax, int rlo, int rhi,           \
                  const double* W, int64_t LW, double* M, int64_t ldM,       \
                  int nthreads) {                                            \
    int nt = nthreads_or_default(nthreads);                                   \
    int cap = 2 * lmax + 1;                                                   \
    _Pragma("omp parallel num_threads(nt)")                                   \
    {                                                                         \
      T* buf0 = (T*)malloc(sizeof(T) * cap);                                  \
      T* buf2 = (T*)malloc(sizeof(T) * cap);                                  \
      _Pragma("omp for schedule(dynamic, 1)")                                 \
      for (int l1 = rlo; l1 <= rhi; ++l1) {                                   \
        for (int l2 = l1; l2 <= lmax; ++l2) {                                 \
          int nmin0, nmin2, n;                                                \
          T xi;                                                               \
          if (kind == PSX_MCM_00) {                                           \
            n = FAMILY(l1, l2, 0, 0, buf0, &nmin0);                           \
            for (int k = 0; k < n; ++k) buf0[k] = buf0[k] * buf0[k];          \
            xi = XI(W, LW, buf0, nmin0, nmin0 + n - 1, l1, l2, 0);            \
          } else if (kind == PSX_MCM_02) {                                    \
            n = FAMILY(l1, l2, 0, 0, buf0, &nmin0);                           \
            FAMILY(l1, l2, -2, 2, buf2, &nmin2);                              \
            for (int k = 0; k < n; ++k) buf0[k] *= buf2[k];                   \
            xi = XI(W, LW, buf0, nmin0, nmin0 + n - 1, l1, l2, 1);            \
          } else {                                                            \
            n = FAMILY(l1, l2, -2, 2, buf2, &nmin2);                          \
            for (int k = 0; k < n; ++k) buf2[k] = buf2[k] * buf2[k];          \
            xi = XI(W, LW, buf2, nmin2, nmin2 + n - 1, l1, l2,                \
                    kind == PSX_MCM_PP ? 1 : 2);                              \
          }                                                                   \
          M[(l1 - lmin) + (int64_t)(l2 - lmin) * ldM] = (double)((T)(2 * l2 + 1) * xi); \
          M[(l2 - lmin) + (int64_t)(l1 - lmin) * ldM] = (double)((T)(2 * l1 + 1) * xi); \
        }                                                                     \
      }                                                                       \
      free(buf0);                                                             \
      free(buf2);                                                             \
    }                                                                         \
    return PSX_OK;                                                            \
  }
DEFINE_MCM(mcm_d, double, wigner_family_d, xi_d)
DEFINE_MCM(mcm_ld, long double, wigner_family_ld, xi_ld)

int psxo_mcm_rows(int kind, int lmin, int lmax, int l1_lo, int l1_hi,
                  const double* W, int64_t LW, double* M, int64_t ldM,
                  int nthreads, int precision) {
  if (kind < PSX_MCM_00 || kind > PSX_MCM_MM) return PSX_ERR_ARGUMENT;
  if (lmin < 0 || lmax < lmin || LW < 1 || ldM < lmax - lmin + 1)
    return PSX_ERR_ARGUMENT;
  if (l1_lo < lmin || l1_hi > lmax || l1_hi < l1_lo) return PSX_ERR_ARGUMENT;
  if (precision == 1)
    return mcm_ld(kind, lmin, lmax, l1_lo, l1_hi, W, LW, M, ldM, nthreads);
  return mcm_d(kind, lmin, lmax, l1_lo, l1_hi, W, LW, M, ldM, nthreads);
}

int psxo_mcm(int kind, int lmin, int lmax, const double* W, int64_t LW,
             double* M, int64_t ldM, int nthreads, int precision) {
  return psxo_mcm_rows(kind, lmin, lmax, lmin, lmax, W, LW, M, ldM, nthreads,
                       precision);
}

/* ------------------------------------------------------------------------
 * Covariance loops, reference src/covariance.jl:92-446.  Spectra, ratios and
 * W come in the reference loop's argument order.
 * ---------------------------------------------------------------------- */
#define XI_TT(Wk, w) xi_d(Wk, LW, w, nmin, nmax, l1, l2, 0)
#define XI_EE(Wk, w) xi_d(Wk, LW, w, nmin, nmax, l1, l2, 1)
#define XI_TE(Wk, w) xi_d(Wk, LW, w, nmin, nmax, l1, l2, 1)

int psxo_cov(int kind, int lmin, int lmax, const double* const* S,
             const double* const* R, const double* const* Wv, int64_t LW,
             double* C, int64_t ldC, int nthreads) {
  return psxo_cov_rows(kind, lmin, lmax, lmin, lmax, S, R, Wv, LW, C, ldC,
                       nthreads);
}

int psxo_cov_rows(int kind, int lmin, int lmax, int l1_lo, int l1_hi,
                  const double* const* S, const double* const* R,
                  const double* const* Wv, int64_t LW, double* C, int64_t ldC,
                  int nthreads) {
  if (kind < PSX_COV_TTTT || kind > PSX_COV_TTEE) return PSX_ERR_ARGUMENT;
  if (lmin < 0 || lmax < lmin || LW < 1 || ldC < lmax - lmin + 1)
    return PSX_ERR_ARGUMENT;
  if (l1_lo < lmin || l1_hi > lmax || l1_hi < l1_lo) return PSX_ERR_ARGUMENT;
  int nt = nthreads_or_default(nthreads);
  int cap = 2 * lmax + 1;
  int domain_error = 0;
#pragma omp parallel num_threads(nt) reduction(| : domain_error)
  {
    double* b0 = (double*)malloc(sizeof(double) * cap);
    double* b2 = (double*)malloc(sizeof(double) * cap);
#pragma omp for schedule(dynamic, 1)
    for (int l1 = l1_lo; l1 <= l1_hi; ++l1) {
      for (int l2 = l1; l2 <= lmax; ++l2) {
        int nmin, nmin2, n = 0;
        double c = 0;
        int nmax;
        const int a = l1, b = l2;
#define SQ(expr) sq_checked((expr), &domain_error)
        switch (kind) {
          case PSX_COV_TTTT: case PSX_COV_EEEE: {
            /* :104-119 / :165-180 */
            int spin2 = kind == PSX_COV_EEEE;
            n = spin2 ? wigner_family_d(l1, l2, -2, 2, b0, &nmin)
                      : wigner_family_d(l1, l2, 0, 0, b0, &nmin);
            for (int k = 0; k < n; ++k) b0[k] = b0[k] * b0[k];
            nmax = nmin + n - 1;
            const double *ip = S[0], *jq = S[1], *iq = S[2], *jp = S[3];
            const double *rip = R[0], *rjq = R[1], *riq = R[2], *rjp = R[3];
            double x[8];
            for (int k = 0; k < 8; ++k)
              x[k] = spin2 ? XI_EE(Wv[k], b0) : XI_TT(Wv[k], b0);
            c = SQ(ip[a] * ip[b] * jq[a] * jq[b]) * x[0] +
                SQ(iq[a] * iq[b] * jp[a] * jp[b]) * x[1] +
                SQ(ip[a] * ip[b]) * x[2] * rjq[a] * rjq[b] +
                SQ(jq[a] * jq[b]) * x[3] * rip[a] * rip[b] +
                SQ(iq[a] * iq[b]) * x[4] * rjp[a] * rjp[b] +
                SQ(jp[a] * jp[b]) * x[5] * riq[a] * riq[b] +
                x[6] * rip[a] * rjq[a] * rip[b] * rjq[b] +
                x[7] * riq[a] * rjp[a] * riq[b] * rjp[b];
            break;
          }
          case PSX_COV_TTTE: {
            /* :219-230 */
            n = wigner_family_d(l1, l2, 0, 0, b0, &nmin);
            for (int k = 0; k < n; ++k) b0[k] = b0[k] * b0[k];
            nmax = nmin + n - 1;
            const double *TTip = S[0], *TTjp = S[1], *TEiq = S[2], *TEjq = S[3];
            const double *rip = R[0], *rjp = R[1];
            c = (SQ(TTip[a] * TTip[b]) * (TEjq[a] + TEjq[b]) * XI_TT(Wv[0], b0) +
                 SQ(TTjp[a] * TTjp[b]) * (TEiq[a] + TEiq[b]) * XI_TT(Wv[1], b0) +
                 (TEjq[a] + TEjq[b]) * XI_TT(Wv[2], b0) * rip[a] * rip[b] +
                 (TEiq[a] + TEiq[b]) * XI_TT(Wv[3], b0) * rjp[a] * rjp[b]) /
                2;
            break;
          }
          case PSX_COV_TETE: {
            /* :276-297 (b2 := f22 * f00, b0 := f00 * f00) */
            n = wigner_family_d(l1, l2, 0, 0, b0, &nmin);
            wigner_family_d(l1, l2, -2, 2, b2, &nmin2);
            for (int k = 0; k < n; ++k) b2[k] *= b0[k];
            for (int k = 0; k < n; ++k) b0[k] *= b0[k];
            nmax = nmin + n - 1;
            const double *TTip = S[0], *EEjq = S[1], *TEiq = S[2], *TEjp = S[3];
            const double *rTTip = R[0], *rPPjq = R[1];
            c = SQ(TTip[a] * TTip[b] * EEjq[a] * EEjq[b]) * XI_TE(Wv[0], b2) +
                0.5 * (TEiq[a] * TEjp[b] + TEjp[a] * TEiq[b]) * XI_TT(Wv[1], b0) +
                SQ(TTip[a] * TTip[b]) * XI_TE(Wv[2], b2) * rPPjq[a] * rPPjq[b] +
                SQ(EEjq[a] * EEjq[b]) * XI_TE(Wv[3], b2) * rTTip[a] * rTTip[b] +
                XI_TE(Wv[4], b2) * rTTip[a] * rTTip[b] * rPPjq[a] * rPPjq[b];
            break;
          }
          case PSX_COV_TEEE: case PSX_COV_TEEE_PLANCK: {
            /* :350-367 (exact: f22*f00) / :387-397 (Planck: f22^2) */
            if (kind == PSX_COV_TEEE) {
              n = wigner_family_d(l1, l2, 0, 0, b0, &nmin);
              wigner_family_d(l1, l2, -2, 2, b2, &nmin2);
              for (int k = 0; k < n; ++k) b2[k] *= b0[k];
            } else {
              n = wigner_family_d(l1, l2, -2, 2, b2, &nmin);
              for (int k = 0; k < n; ++k) b2[k] = b2[k] * b2[k];
            }
            nmax = nmin + n - 1;
            const double *EEjq = S[0], *EEjp = S[1], *TEip = S[2], *TEiq = S[3];
            const double *rjq = R[0], *rjp = R[1];
            c = (SQ(EEjq[a] * EEjq[b]) * (TEip[a] + TEip[b]) * XI_EE(Wv[0], b2) +
                 SQ(EEjp[a] * EEjp[b]) * (TEiq[a] + TEiq[b]) * XI_EE(Wv[1], b2) +
                 (TEip[a] + TEip[b]) * XI_EE(Wv[2], b2) * rjq[a] * rjq[b] +
                 (TEiq[a] + TEiq[b]) * XI_EE(Wv[3], b2) * rjp[a] * rjp[b]) /
                2;
            break;
          }
          case PSX_COV_TTEE: {
            /* :433-441 */
            n = wigner_family_d(l1, l2, 0, 0, b0, &nmin);
            for (int k = 0; k < n; ++k) b0[k] = b0[k] * b0[k];
            nmax = nmin + n - 1;
            const double *TEip = S[0], *TEiq = S[1], *TEjq = S[2], *TEjp = S[3];
            c = ((TEip[a] * TEjq[b] + TEjq[a] * TEip[b]) * XI_TT(Wv[0], b0) +
                 (TEiq[a] * TEjp[b] + TEjp[a] * TEiq[b]) * XI_TT(Wv[1], b0)) /
                2;
            break;
          }
        }
#undef SQ
        C[(l1 - lmin) + (int64_t)(l2 - lmin) * ldC] = c;
        C[(l2 - lmin) + (int64_t)(l1 - lmin) * ldC] = c;
      }
    }
    free(b0);
    free(b2);
  }
  return domain_error ? PSX_ERR_DOMAIN : PSX_OK;
}

static int cov_arity(int kind, int* nW, int* nspectra, int* nratios) {
  static const int tab[7][3] = {{8, 4, 4}, {8, 4, 4}, {4, 4, 2}, {5, 4, 2},
                                {4, 4, 2}, {4, 4, 2}, {2, 4, 0}};
  if (kind < 0 || kind > 6) return PSX_ERR_ARGUMENT;
  if (nW) *nW = tab[kind][0];
  if (nspectra) *nspectra = tab[kind][1];
  if (nratios) *nratios = tab[kind][2];
  return PSX_OK;
}

/* ------------------------------------------------------------------------
 * Dense LU with partial pivoting (LAPACK getrf semantics: pivot = first row
 * of maximal |a|), reference use: src/covariance.jl:11-12 (lu + rdiv!) and
 * src/blockspectralmatrix.jl:124-129 (lu + \).
 * ---------------------------------------------------------------------- */
static int lu_factor(int n, double* A, int* piv) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    double mx = fabs(A[k + (int64_t)k * n]);
    for (int i = k + 1; i < n; ++i) {
      double v = fabs(A[i + (int64_t)k * n]);
      if (v > mx) { mx = v; p = i; }
    }
    piv[k] = p;
    if (mx == 0) return PSX_ERR_SINGULAR;
    if (p != k)
      for (int j = 0; j < n; ++j) {
        double t = A[k + (int64_t)j * n];
        A[k + (int64_t)j * n] = A[p + (int64_t)j * n];
        A[p + (int64_t)j * n] = t;
      }
    double inv = 1.0 / A[k + (int64_t)k * n];
    for (int i = k + 1; i < n; ++i) A[i + (int64_t)k * n] *= inv;
    for (int j = k + 1; j < n; ++j) {
      double akj = A[k + (int64_t)j * n];
      if (akj != 0)
        for (int i = k + 1; i < n; ++i)
          A[i + (int64_t)j * n] -= A[i + (int64_t)k * n] * akj;
    }
  }
  return PSX_OK;
}

/* Solve (LU) X = B in place for nrhs columns. */
static void lu_solve_cols(int n, const double* LU, const int* piv, double* B,
                          int nrhs) {
  for (int r = 0; r < nrhs; ++r) {
    double* b = B + (int64_t)r * n;
    for (int k = 0; k < n; ++k)
      if (piv[k] != k) { double t = b[k]; b[k] = b[piv[k]]; b[piv[k]] = t; }
    for (int j = 0; j < n; ++j) {
      double bj = b[j];
      if (bj != 0)
        for (int i = j + 1; i < n; ++i) b[i] -= LU[i + (int64_t)j * n] * bj;
    }
    for (int j = n - 1; j >= 0; --j) {
      b[j] /= LU[j + (int64_t)j * n];
      double bj = b[j];
      if (bj != 0)
        for (int i = 0; i < j; ++i) b[i] -= LU[i + (int64_t)j * n] * bj;
    }
  }
}

int psxo_lu_solve(int n, const double* A, double* B, int nrhs) {
  if (n < 1 || nrhs < 0) return PSX_ERR_ARGUMENT;
  double* LU = (double*)malloc(sizeof(double) * (size_t)n * n);
  int* piv = (int*)malloc(sizeof(int) * n);
  memcpy(LU, A, sizeof(double) * (size_t)n * n);
  int st = lu_factor(n, LU, piv);
  if (st == PSX_OK) lu_solve_cols(n, LU, piv, B, nrhs);
  free(LU);
  free(piv);
  return st;
}

/* decouple_covmat, reference src/covariance.jl:8-14:
 *   C = copy(Y); C' <- C' / B1' (i.e. C <- B1^{-1} C); C <- C / B2'
 *   (i.e. C <- C B2^{-T}, equivalently C^T <- B2^{-1} C^T). */
int psxo_decouple_covmat(int n, const double* Y, const double* B1,
                         const double* B2, double* out) {
  if (n < 1) return PSX_ERR_ARGUMENT;
  size_t nn = (size_t)n * n;
  double* T = (double*)malloc(sizeof(double) * nn);
  memcpy(T, Y, sizeof(double) * nn);
  int st = psxo_lu_solve(n, B1, T, n); /* T = B1^{-1} Y */
  if (st == PSX_OK) {
    for (int i = 0; i < n; ++i) /* out = T^T */
      for (int j = 0; j < n; ++j) out[j + (int64_t)i * n] = T[i + (int64_t)j * n];
    st = psxo_lu_solve(n, B2, out, n); /* out = B2^{-1} T^T */
    if (st == PSX_OK) {
      memcpy(T, out, sizeof(double) * nn);
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) out[j + (int64_t)i * n] = T[i + (int64_t)j * n];
    }
  }
  free(T);
  return st;
}

/* ------------------------------------------------------------------------
 * QuickPol Ξ, reference src/beam.jl:17-28 (Ξsum), :57-61 (specrowrange) and
 * :72-101 (quickpolΞ!).  The matrix is a BandedMatrix (lower bandwidth bl,
 * upper bu) in band storage: element (row i = l'', column j = l), 0-based,
 * at band[(bu + i - j) + j * (bl + bu + 1)].  For every row l'' in 2..lmax
 * and column l in max(2, l'' - bl)..min(lmax, l'' + bu):
 *   Ξ[l'', l] = sum_{l'} W[l'] f1(l') f2(l'),
 *   f1 = WignerF(l, l'', -s1, -nu1) = (l' l l''; s1+nu1, -s1, -nu1),
 *   f2 = WignerF(l, l'', -s2, -nu2),
 * l' over the intersection of the two families' ranges.  Then the whole band
 * (the entries just written and all others) is multiplied by
 * (-1)^(s1 + s2 + nu1 + nu2) (beam.jl:98-99), so entries the loop does not
 * write (rows or columns 0 and 1) keep their input value times that sign.
 * The reference indexes W at l' without bounds checks (@inbounds); here terms
 * with l' >= LW are zero, identical whenever LW > 2 lmax.
 * ---------------------------------------------------------------------- */
int psxo_quickpol_xi(int nu1, int nu2, int s1, int s2, int lmax, int bl, int bu,
                     const double* W, int64_t LW, double* band, int nthreads) {
  if (lmax < 0 || bl < 0 || bu < 0 || LW < 0) return PSX_ERR_ARGUMENT;
  const int64_t ldb = (int64_t)bl + bu + 1;
  const double sgn = ((s1 + s2 + nu1 + nu2) % 2 == 0) ? 1.0 : -1.0;
  int nt = nthreads_or_default(nthreads);
  int status = PSX_OK;
#pragma omp parallel num_threads(nt)
  {
    double* w1 = (double*)malloc(sizeof(double) * (2 * (size_t)lmax + 2));
    double* w2 = (double*)malloc(sizeof(double) * (2 * (size_t)lmax + 2));
    if (!w1 || !w2) {
#pragma omp atomic write
      status = PSX_ERR_BUDGET;
    }
#pragma omp for schedule(dynamic)
    for (int lpp = 2; lpp <= lmax; ++lpp) {
      if (!w1 || !w2) continue;
      const int lo = lpp - bl > 2 ? lpp - bl : 2;
      const int hi = lpp + bu < lmax ? lpp + bu : lmax;
      for (int l = lo; l <= hi; ++l) {
        int n1min, n2min;
        const int n1 = wigner_family_d(l, lpp, -s1, -nu1, w1, &n1min);
        const int n2 = wigner_family_d(l, lpp, -s2, -nu2, w2, &n2min);
        const int start = n1min > n2min ? n1min : n2min;
        int64_t stop = l + lpp;                 /* both families end at l + l'' */
        if (n1 <= 0 || n2 <= 0) stop = start - 1;
        if (stop > LW - 1) stop = LW - 1;
        double xi = 0.0;
        for (int64_t lp = start; lp <= stop; ++lp)
          xi += W[lp] * w1[lp - n1min] * w2[lp - n2min];
        band[(bu + lpp - l) + (int64_t)l * ldb] = xi;
      }
    }
    free(w1);
    free(w2);
  }
  /* 𝚵 .*= sgn over the stored band */
  const int n = lmax + 1;
  for (int j = 0; j < n; ++j)
    for (int d = -bl; d <= bu; ++d) {
      const int i = j - d;                       /* row with j - i = d */
      if (i < 0 || i >= n) continue;
      band[(bu - d) + (int64_t)j * ldb] *= sgn;
    }
  return status;
}
