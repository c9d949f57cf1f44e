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
// normalisation sums are accumulated on the fly, so nothing is stored.  The
// factor of A(j+1) that depends on j, d and M1 only is the same for every
// lane: the wavefront computes 64 steps of it at a time into LDS (one step
// per lane) together with those steps' W values, which leaves one reciprocal
// square root per lane and step and no global load in the recurrence:
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
  int nd, ntask;    // diagonals, wavefront tasks (nd x row blocks)
};

// Launch order of the wavefront tasks.  Task t = (row block t / nd, diagonal
// t % nd) runs families of about 2 (lmax - 64 (t / nd)) steps, longest first.
// All wavefronts are resident at once (~2 per SIMD), and consecutive
// workgroups go to different SIMDs, so the tasks are taken in layers of
// kQpLayer with every odd layer reversed: the SIMD that holds one of the
// longest families of the first layer also gets one of the shortest of the
// second.
#ifndef PSX_QP_PRIO
#define PSX_QP_PRIO 1
#endif
#ifndef PSX_QP_LAYER
#define PSX_QP_LAYER 1024
#endif
constexpr int kQpLayer = PSX_QP_LAYER;

__device__ __forceinline__ int qp_task(int b, int ntask) {
  const int L = b / kQpLayer, k = b - L * kQpLayer;
  if (!(L & 1)) return b;
  const int size = min(kQpLayer, ntask - L * kQpLayer);
  return L * kQpLayer + (size - 1 - k);
}

constexpr double kQpBig = 1e150;
constexpr double kQpTiny = 1e-150;
constexpr int kQpSteps = 64;      // recurrence steps per LDS table refill (= lanes)

struct QpState {
  double f = 1.0, p = 0.0, A = 0.0, S = 0.0;
};

// f(j+1) from f(j), f(j-1) for j >= 1.  au = sqrt(((j+1)^2-d^2)((j+1)^2-M1^2))
// and rj = 1/(j au) are the wavefront-uniform parts of A(j+1) = au rv and of
// 1/(j A(j+1)) = rj r, with rv = sqrt((l+l''+1)^2-(j+1)^2) and r = 1/rv per
// lane; s.A holds A(j) on entry and A(j+1) on return.  jp = j + 1,
// tj = 2j + 1, kj = (2j+1) j (j+1).
__device__ __forceinline__ void qp_adv(QpState& s, double au, double rj, double rv, double r,
                                       double jp, double tj, double kj, double c, double e) {
  const double Y = fma(kj, e, -tj * c);        // -(2j+1)[c - j(j+1) e]
  const double fn = -fma(Y, s.f, jp * s.A * s.p) * (rj * r);
  s.p = s.f;
  s.f = fn;
  s.A = au * rv;
}

// The main-loop step.  With rj(j) = 1/(j au(j)) folded in, the recurrence
// reads f(j+1) = -[Yr f(j) + C(j) f(j-1)] r, Yr = Y(j) rj(j), and the carried
// coefficient C(j) = (j+1) A(j) rj(j) (held in s.A) advances as
// C(j+1) = g(j) rv with the uniform g(j) = (j+2) au(j) rj(j+1).
__device__ __forceinline__ void qp_step(QpState& s, double Yr, double g, double rv, double r) {
  const double fn = -fma(Yr, s.f, s.A * s.p) * r;
  s.p = s.f;
  s.f = fn;
  s.A = g * rv;
}

// 1/sqrt(v) for finite v > 0: the hardware estimate r0 (about 23 correct
// bits) and one Newton-Raphson step carried to second order,
// h = 1 - v r0^2, r = r0 (1 + h/2 + 3h^2/8), which leaves about one ulp.
__device__ __forceinline__ double qp_rsqrt(double v) {
  const double r0 = __builtin_amdgcn_rsq(v);
  const double h = fma(-v * r0, r0, 1.0);
  return fma(r0 * h, fma(0.375, h, 0.5), r0);
}

// The step out of j = 0 (nmin = 0: l == l'', M1 == 0), where the recurrence
// divided by j gives A(1) f(1) + (M3 - M2) f(0) = 0.
__device__ __forceinline__ void qp_adv0(QpState& s, double au, double rv, double e) {
  const double An = au * rv;
  s.p = s.f;
  s.f = -e / An;
  s.A = An;
}

__device__ __forceinline__ void qp_rescale(QpState& s, double& D) {
  if (fmax(fabs(s.f), fabs(s.p)) > kQpBig) {
    s.f *= kQpTiny;
    s.p *= kQpTiny;
    s.S *= kQpTiny * kQpTiny;
    D *= kQpTiny;
  }
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}

__global__ void __launch_bounds__(64) quickpol_xi_kernel(QpArgs A) {
  // per step and family (a, b): tj rj, kj e rj, g (qp_step); then W, tj =
  // 2j + 1 and (j + 1)^2
  __shared__ double tab[9][kQpSteps];
  const int task = qp_task((int)blockIdx.x, A.ntask);
  const int d = task % A.nd - A.bl;                                       // l - l''
  const int lpp0 = A.lmax - (task / A.nd) * 64;                           // lane 0's l''
  const int lane = (int)threadIdx.x;
  const int lpp = lpp0 - lane;                                            // l'', descending
  const int l = lpp + d;
  if (lpp0 < 2 || lpp0 + d < 2) {
    // no lane of this wavefront has l'' >= 2 and l >= 2 (lane 0 has the largest)
    if (lpp >= 0 && l >= 0 && l <= A.lmax) A.band[(A.bu - d) + (int64_t)l * A.ldb] *= A.sgn;
    return;
  }
  const bool inmat = lpp >= 0 && l >= 0 && l <= A.lmax;
  double* cell = A.band + (A.bu - d) + (int64_t)(inmat ? l : 0) * A.ldb;
  bool live = false;
  if (inmat) {
    if (lpp < 2 || l < 2) {
      *cell *= A.sgn;                             // not written by the loop, only signed
    } else if (abs(A.a.M2) > l || abs(A.a.M3) > lpp || abs(A.b.M2) > l || abs(A.b.M3) > lpp) {
      *cell = 0.0 * A.sgn;                        // no valid 3j symbol: the family is zero
    } else {
      live = true;
    }
  }
  const int ad = d < 0 ? -d : d;
  const int na = max(ad, abs(A.a.M1)), nb = max(ad, abs(A.b.M1));
  const int jc = max(na, nb);
  const int jend = 2 * lpp0 + d;                  // largest l + l'' of the wavefront
#if PSX_QP_PRIO
  // issue priority by chain length: where a SIMD holds a long and a short
  // wavefront, the long one (which sets the kernel's time) issues first
  if (4 * jend >= 3 * 2 * A.lmax) __builtin_amdgcn_s_setprio(3);
  else if (2 * jend >= 2 * A.lmax) __builtin_amdgcn_s_setprio(2);
  else if (4 * jend >= 2 * A.lmax) __builtin_amdgcn_s_setprio(1);
#endif
  // lanes without a cell run the recurrence of lane 0's cell and discard it
  const int nmax = live ? l + lpp : jend;
  const int ell = live ? l : lpp0 + d, ellpp = live ? lpp : lpp0;
  const int nmin_end = wave_min(live ? nmax : jend);   // every lane is active below it
  const double L2 = (double)ell * (ell + 1) - (double)ellpp * (ellpp + 1);
  const double ca = A.a.M1 * L2, cb = A.b.M1 * L2;
  const double ea = A.a.M3 - A.a.M2, eb = A.b.M3 - A.b.M2;
  const double ma2 = (double)A.a.M1 * A.a.M1, mb2 = (double)A.b.M1 * A.b.M1;
  const double s2 = (double)(nmax + 1) * (nmax + 1);
  const double d2 = (double)d * d;
  QpState fa, fb;
  double D = 0.0;

  // prologue: the family with the lower start runs alone up to jc (at most
  // a few steps, |M1a| - |M1b| apart), then both stand at j = jc
  {
    QpState& f1 = na < nb ? fa : fb;
    const double m2 = na < nb ? ma2 : mb2;
    const double c1 = na < nb ? ca : cb, e1 = na < nb ? ea : eb;
    for (int j = min(na, nb); j < jc; ++j) {
      const double jp = j + 1.0, tj = 2.0 * j + 1.0;
      f1.S = fma(tj * f1.f, f1.f, f1.S);
      const double v = s2 - jp * jp;
      const double rv = sqrt(v);
      const double au = sqrt((jp * jp - d2) * (jp * jp - m2));
      if (j == 0) qp_adv0(f1, au, rv, e1);
      else qp_adv(f1, au, 1.0 / ((double)j * au), rv, 1.0 / rv, jp, tj, tj * j * jp, c1, e1);
    }
  }
  int j1 = jc;
  if (jc == 0) {
    // both families start at 0 (l == l'', M1 = 0 for both): peel the j = 0 step
    fa.S = fa.f * fa.f;
    fb.S = fb.f * fb.f;
    if (A.LW > 0) D = A.W[0] * fa.f * fb.f;
    const double rv = sqrt(s2 - 1.0);
    qp_adv0(fa, sqrt((1.0 - d2) * (1.0 - ma2)), rv, ea);
    qp_adv0(fb, sqrt((1.0 - d2) * (1.0 - mb2)), rv, eb);
    j1 = 1;
  }
  // the main loop carries C(j1) = (j1+1) A(j1) rj(j1) in place of A(j1)
  {
    const double q = ((double)j1 + 1.0) * ((double)j1 + 1.0);
    fa.A *= ((double)j1 + 1.0) / ((double)j1 * sqrt((q - d2) * (q - ma2)));
    fb.A *= ((double)j1 + 1.0) / ((double)j1 * sqrt((q - d2) * (q - mb2)));
  }

  // main part: both families step together, uniform j; steps below
  // nmin_end need no per-lane test
  for (int jb = j1; jb <= jend; jb += kQpSteps) {
    {
      const int j = jb + lane;
      const double jp = (double)j + 1.0, jq = (double)j + 2.0;
      const double jp2 = jp * jp, jq2 = jq * jq;
      const double tj = 2.0 * (double)j + 1.0, kj = tj * (double)j * jp;
      double yta = 0.0, yka = 0.0, ytb = 0.0, ykb = 0.0, ga = 0.0, gb = 0.0;
      if (j < jend) {
        // au(j) = sqrt(((j+1)^2 - d^2)((j+1)^2 - M1^2)) > 0 for j >= j1
        const double aa = sqrt((jp2 - d2) * (jp2 - ma2)), ab = sqrt((jp2 - d2) * (jp2 - mb2));
        const double aa1 = sqrt((jq2 - d2) * (jq2 - ma2));
        const double ab1 = sqrt((jq2 - d2) * (jq2 - mb2));
        const double ra = 1.0 / ((double)j * aa), rb = 1.0 / ((double)j * ab);
        yta = tj * ra;
        yka = kj * ea * ra;
        ytb = tj * rb;
        ykb = kj * eb * rb;
        ga = jq * aa / (jp * aa1);
        gb = jq * ab / (jp * ab1);
      }
      const double w = j <= jend && j < A.LW ? A.W[j] : 0.0;
      __syncthreads();
      tab[0][lane] = yta;
      tab[1][lane] = yka;
      tab[2][lane] = ga;
      tab[3][lane] = ytb;
      tab[4][lane] = ykb;
      tab[5][lane] = gb;
      tab[6][lane] = w;
      tab[7][lane] = tj;
      tab[8][lane] = jp2;
      __syncthreads();
    }
    const int tend = min(kQpSteps, jend - jb + 1);
    const int tm = max(0, min(tend, nmin_end - jb));   // steps with j < nmin_end
    for (int t = 0; t < tm; ++t) {
      const double tj = tab[7][t];
      fa.S = fma(tj * fa.f, fa.f, fa.S);
      fb.S = fma(tj * fb.f, fb.f, fb.S);
      D = fma(tab[6][t] * fa.f, fb.f, D);      // W is zero beyond LW
      const double v = s2 - tab[8][t];          // (l+l''+1)^2 - (j+1)^2 > 0
      const double r = qp_rsqrt(v);
      const double rv = v * r;
      qp_step(fa, fma(-tab[0][t], ca, tab[1][t]), tab[2][t], rv, r);
      qp_step(fb, fma(-tab[3][t], cb, tab[4][t]), tab[5][t], rv, r);
    }
    for (int t = tm; t < tend; ++t) {          // the last steps, lane by lane
      const int j = jb + t;
      if (j > nmax) break;
      const double tj = tab[7][t];
      fa.S = fma(tj * fa.f, fa.f, fa.S);
      fb.S = fma(tj * fb.f, fb.f, fb.S);
      D = fma(tab[6][t] * fa.f, fb.f, D);
      if (j == nmax) break;
      const double v = s2 - tab[8][t];
      const double r = qp_rsqrt(v);
      const double rv = v * r;
      qp_step(fa, fma(-tab[0][t], ca, tab[1][t]), tab[2][t], rv, r);
      qp_step(fb, fma(-tab[3][t], cb, tab[4][t]), tab[5][t], rv, r);
    }
    qp_rescale(fa, D);
    qp_rescale(fb, D);
  }
  if (!live) return;
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
  // the forward recurrence and its once-per-64-steps overflow check are
  // sized for scan-pattern spins
  for (int v : sp)
    PSX_REQUIRE(v >= -8 && v <= 8, PSX_ERR_ARGUMENT, "spins must be within -8..8");
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
  A.nd = bl + bu + 1;
  A.ntask = A.nd * ((lmax + 1 + 63) / 64);
  const dim3 grid((unsigned)A.ntask);
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
    const size_t wb = sizeof(double) * (LW > 0 ? LW : 1),
                 bb = sizeof(double) * ldb * (size_t)(lmax + 1);
    XferBuf x(c, xfer_bytes({wb, bb}));
    double* dW = x.take(wb);
    double* dB = x.take(bb);
    if (LW > 0) h2d_matrix(c, dW, LW, W, LW, LW, 1);
    h2d_matrix(c, dB, ldb, band, ldb, ldb, lmax + 1);
    quickpol_xi_dev(c, nu1, nu2, s1, s2, lmax, bl, bu, dW, LW, dB);
    d2h_matrix(c, band, ldb, dB, ldb, ldb, lmax + 1);
    PSX_HIP(hipStreamSynchronize(c->stream));
  });
}

}  // extern "C"
