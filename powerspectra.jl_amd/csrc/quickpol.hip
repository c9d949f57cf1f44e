// QuickPol Ξ matrix (reference src/beam.jl:17-28 Ξsum, :57-61 specrowrange,
// :72-101 quickpolΞ!): for every row l'' >= 2 of a banded matrix and every
// column l >= 2 inside the band,
//   Ξ[l'', l] = (-1)^(s1+s2+nu1+nu2) sum_{l'} W[l'] f1(l') f2(l'),
//   f1(l') = (l' l l''; s1+nu1, -s1, -nu1),  f2(l') = (l' l l''; s2+nu2, -s2, -nu2),
// each family normalised by sum (2l'+1) f^2 = 1 with f(l+l'') of sign
// (-1)^(l - l'' + M1) (M1 = s+nu).  Every other stored band entry is
// multiplied by the same sign (the reference's `𝚵 .*= sgn`).
//
// One lane per band cell, one wavefront per (diagonal d = l - l'', 64
// consecutive l''): the two families of every lane start at the same
// j = max(|d|, |M1|), so the recurrence index, the W load and the family
// start are uniform across the wavefront and only the end (l + l'') differs.
// Each family runs the Schulten-Gordon three-term recurrence in j forward
// from its lowest value (f(nmin) = 1), and the sum_j W f1 f2 and both
// normalisation sums are accumulated on the fly, so nothing is stored:
//   j A(j+1) f(j+1) + Y(j) f(j) + (j+1) A(j) f(j-1) = 0,
//   A(j) = sqrt((j^2 - d^2)((l+l''+1)^2 - j^2)(j^2 - M1^2)),
//   Y(j) = -(2j+1)[M1 (l(l+1) - l''(l''+1)) - j(j+1)(M3 - M2)].
// For the small spins of a scan pattern (|s|, |nu| <= a few) the classically
// forbidden zones at both ends of a family are a few terms wide, where the
// forward recurrence stays within ~1e-14 of the family maximum.
#include "psx_internal.h"

#include <cmath>

namespace psx {

namespace {

struct QpFam {
  int M1, M2, M3;   // (j l l''; M1 M2 M3), M1 = -M2 - M3
};

struct QpArgs {
  const double* W;
  int64_t LW;
  double* band;     // (bl + bu + 1) x (lmax + 1), column-major
  int64_t ldb;
  int lmax, bl, bu;
  QpFam a, b;
  double sgn;
};

constexpr double kQpBig = 1e150;
constexpr double kQpTiny = 1e-150;

struct QpState {
  double f = 1.0, p = 0.0, A = 0.0, S = 0.0;
};

// f(j+1) from f(j), f(j-1); A holds A(j) on entry and A(j+1) on return.
__device__ __forceinline__ void qp_step(QpState& s, int j, double jp2, double t, double rv,
                                        double m2, double c, double e, double& D) {
  const double jp = (double)j + 1.0;
  const double An = sqrt(t * (jp2 - m2)) * rv;
  double fn;
  if (j == 0) {
    // nmin = 0 (l == l'', M1 == 0): the recurrence divided by j gives
    // A(1) f(1) + (M3 - M2) f(0) = 0
    fn = -e / An;
  } else {
    const double Y = -(2.0 * j + 1.0) * (c - (double)j * jp * e);
    fn = -(Y * s.f + jp * s.A * s.p) / ((double)j * An);
  }
  s.p = s.f;
  s.f = fn;
  s.A = An;
  if (fabs(fn) > kQpBig) {
    s.f *= kQpTiny;
    s.p *= kQpTiny;
    s.S *= kQpTiny * kQpTiny;
    D *= kQpTiny;
  }
}

__global__ void __launch_bounds__(64) quickpol_xi_kernel(QpArgs A) {
  const int d = (int)blockIdx.x - A.bl;                              // l - l''
  const int lpp = A.lmax - ((int)blockIdx.y * 64 + (int)threadIdx.x);  // l'', descending
  const int l = lpp + d;
  if (lpp < 0 || l < 0 || l > A.lmax) return;
  double* cell = A.band + (A.bu - d) + (int64_t)l * A.ldb;
  if (lpp < 2 || l < 2) {        // not written by the loop, only signed
    *cell *= A.sgn;
    return;
  }
  const int nmax = l + lpp;
  const int ad = d < 0 ? -d : d;
  const int na = max(ad, abs(A.a.M1)), nb = max(ad, abs(A.b.M1));
  const bool va = abs(A.a.M2) <= l && abs(A.a.M3) <= lpp;
  const bool vb = abs(A.b.M2) <= l && abs(A.b.M3) <= lpp;
  if (!va || !vb) {              // no valid 3j symbol: the family is zero
    *cell = 0.0 * A.sgn;
    return;
  }
  const double L2 = (double)l * (l + 1) - (double)lpp * (lpp + 1);
  const double ca = A.a.M1 * L2, cb = A.b.M1 * L2;
  const double ea = A.a.M3 - A.a.M2, eb = A.b.M3 - A.b.M2;
  const double ma2 = (double)A.a.M1 * A.a.M1, mb2 = (double)A.b.M1 * A.b.M1;
  const double s2 = (double)(nmax + 1) * (nmax + 1);
  const double d2 = (double)d * d;
  QpState fa, fb;
  double D = 0.0;
  const int j0 = min(na, nb), jc = max(na, nb);
  const int64_t wend = A.LW;
  for (int j = j0;; ++j) {
    const double tj = 2.0 * j + 1.0;
    if (j >= na) fa.S = fma(tj * fa.f, fa.f, fa.S);
    if (j >= nb) fb.S = fma(tj * fb.f, fb.f, fb.S);
    if (j >= jc && j < wend) D = fma(A.W[j] * fa.f, fb.f, D);
    if (j == nmax) break;
    const double jp = (double)j + 1.0;
    const double jp2 = jp * jp;
    const double rv = sqrt(s2 - jp2);
    const double t = jp2 - d2;
    if (j >= na) qp_step(fa, j, jp2, t, rv, ma2, ca, ea, D);
    if (j >= nb) qp_step(fb, j, jp2, t, rv, mb2, cb, eb, D);
  }
  // normalise, fixing the sign of f(l + l'') to (-1)^(l - l'' + M1)
  double ka = 1.0 / sqrt(fa.S), kb = 1.0 / sqrt(fb.S);
  if ((fa.f < 0.0) != (((l - lpp + A.a.M1) & 1) != 0)) ka = -ka;
  if ((fb.f < 0.0) != (((l - lpp + A.b.M1) & 1) != 0)) kb = -kb;
  *cell = A.sgn * (D * ka * kb);
}

void check_qp(int nu1, int nu2, int s1, int s2, int lmax, int bl, int bu, int64_t LW) {
  PSX_REQUIRE(lmax >= 0 && lmax <= 65535, PSX_ERR_ARGUMENT, "lmax must be 0..65535");
  PSX_REQUIRE(bl >= 0 && bu >= 0 && bl <= lmax && bu <= lmax, PSX_ERR_ARGUMENT,
              "bandwidths must be 0..lmax");
  PSX_REQUIRE(LW >= 0, PSX_ERR_ARGUMENT, "W length must be >= 0");
  const int sp[4] = {nu1, nu2, s1, s2};
  for (int v : sp)
    PSX_REQUIRE(v >= -64 && v <= 64, PSX_ERR_ARGUMENT, "spins must be within -64..64");
}

void quickpol_xi_dev(Ctx* c, int nu1, int nu2, int s1, int s2, int lmax, int bl, int bu,
                     const double* dW, int64_t LW, double* dband) {
  QpArgs A;
  A.W = dW;
  A.LW = LW;
  A.band = dband;
  A.ldb = (int64_t)bl + bu + 1;
  A.lmax = lmax;
  A.bl = bl;
  A.bu = bu;
  // WignerF(T, l, l'', -s, -nu): (j l l''; s+nu, -s, -nu)
  A.a = QpFam{s1 + nu1, -s1, -nu1};
  A.b = QpFam{s2 + nu2, -s2, -nu2};
  A.sgn = ((s1 + s2 + nu1 + nu2) & 1) ? -1.0 : 1.0;
  const dim3 grid((unsigned)(bl + bu + 1), (unsigned)((lmax + 1 + 63) / 64));
  TimedLaunch tl(c, "psx_quickpol_xi");
  hipLaunchKernelGGL(quickpol_xi_kernel, grid, dim3(64), 0, c->stream, A);
  PSX_HIP(hipGetLastError());
}

}  // namespace

}  // namespace psx

using namespace psx;

extern "C" {

int psx_quickpol_xi_d(psx_ctx* h, int nu1, int nu2, int s1, int s2, int lmax, int bl, int bu,
                      const double* dW, int64_t LW, double* dband) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  return guarded(c, [&] {
    check_qp(nu1, nu2, s1, s2, lmax, bl, bu, LW);
    PSX_REQUIRE(dband && (dW || LW == 0), PSX_ERR_ARGUMENT, "null pointer");
    PSX_HIP(hipSetDevice(c->device));
    quickpol_xi_dev(c, nu1, nu2, s1, s2, lmax, bl, bu, dW, LW, dband);
    PSX_HIP(hipStreamSynchronize(c->stream));
  });
}

int psx_quickpol_xi(psx_ctx* h, int nu1, int nu2, int s1, int s2, int lmax, int bl, int bu,
                    const double* W, int64_t LW, double* band) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  return guarded(c, [&] {
    check_qp(nu1, nu2, s1, s2, lmax, bl, bu, LW);
    PSX_REQUIRE(band && (W || LW == 0), PSX_ERR_ARGUMENT, "null pointer");
    PSX_HIP(hipSetDevice(c->device));
    const int64_t ldb = (int64_t)bl + bu + 1;
    double* dW = (double*)ctx_alloc(c, sizeof(double) * (LW > 0 ? LW : 1));
    double* dB = nullptr;
    try {
      dB = (double*)ctx_alloc(c, sizeof(double) * ldb * (lmax + 1));
      if (LW > 0) h2d_matrix(c, dW, LW, W, LW, LW, 1);
      h2d_matrix(c, dB, ldb, band, ldb, ldb, lmax + 1);
      quickpol_xi_dev(c, nu1, nu2, s1, s2, lmax, bl, bu, dW, LW, dB);
      d2h_matrix(c, band, ldb, dB, ldb, ldb, lmax + 1);
      PSX_HIP(hipStreamSynchronize(c->stream));
    } catch (...) {
      (void)hipStreamSynchronize(c->stream);
      ctx_free(c, dW);
      if (dB) ctx_free(c, dB);
      throw;
    }
    ctx_free(c, dW);
    ctx_free(c, dB);
  });
}

}  // extern "C"
