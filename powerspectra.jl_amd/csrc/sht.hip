// Spherical-harmonic transforms of ring-ordered HEALPix maps on the GPU:
// the mask -> a_lm -> W_l front end of mcm(spec, mask1, mask2) and of the
// covariance workspace (reference src/modecoupling.jl:197,250-256,
// src/workspace.jl:141-213, which call Healpix.jl's map2alm / alm2map).
//
// Analysis of one map (spin 0):
//   1. ring Fourier coefficients.  The 2 nside + 1 equatorial rings all have
//      4 nside pixels: one batched real-to-complex rocFFT.  The 2 (nside - 1)
//      polar rings (4i pixels) go through a direct DFT kernel.  Each ring keeps
//      its Hermitian half X[0 .. n/2].
//   2. fold: F_k(m) = X_k[m mod n] (conjugated above n/2) e^{-i m phi0_k},
//      combined over the north/south ring pair into F+ = F_N + F_S and
//      F- = F_N - F_S (lambda_lm(-z) = (-1)^{l+m} lambda_lm(z)).
//   3. Legendre sums a_lm = (4 pi / Npix) sum_pairs lambda_lm(z) F+-(m), one
//      workgroup per (m, chunk of 1024 ring pairs), the sum over rings
//      reduced through LDS in rounds of 8 l, partial sums per chunk added by
//      a second kernel.
// Synthesis runs the same steps backwards (one thread per ring pair
// accumulating over l, no reduction), and map2alm refines with `niter` Jacobi
// iterations a += A(map - S(a)) as Healpix.jl does.
//
// lambda_lm is the orthonormal associated Legendre function (Y_lm = lambda_lm
// e^{i m phi}, Condon-Shortley phase) from the three-term recurrence in l,
// started at lambda_mm = (-1)^m sqrt((2m+1)/(4 pi) prod_k (2k-1)/(2k)) sin^m.
// That start underflows double range long before it stops mattering, so it
// is carried as mantissa x 2^(600 s), s <= 0, mantissa <= 2^600 (rescaled
// whenever the recurrence grows past it).  A term contributes its true value
// mantissa x 2^(600 s) when s >= -1; below that it is under 2^-600.
#include "psx_internal.h"

#include <rocfft/rocfft.h>

#include <algorithm>
#include <cmath>
#include <functional>
#include <map>
#include <mutex>
#include <vector>

namespace psx {

namespace {

constexpr int kMaxPolarRing = 8192;         // pixels of the longest polar ring (nside <= 2048)
constexpr int kLChunk = 8;                  // l values per LDS reduction round
#ifndef PSX_SHT_PAIRS
#define PSX_SHT_PAIRS 4
#endif
constexpr int kPairsPerThread = PSX_SHT_PAIRS;   // ring pairs per analysis thread
#ifndef PSX_SYN_PAIRS
#define PSX_SYN_PAIRS 4
#endif
constexpr int kSynPairs = PSX_SYN_PAIRS;     // ring pairs per synthesis thread
// The spin-2 synthesis runs 2 ring pairs per thread at 5 waves per SIMD
// (96 VGPRs, 28 bytes spilled outside the l loop): more waves cover its four
// dependent recurrence chains better than more pairs per thread (4 pairs at
// 3 waves took 1.8x as long at nside 1024, lmax 3071).
#ifndef PSX_SYN2_PAIRS
#define PSX_SYN2_PAIRS 2
#endif
#ifndef PSX_SYN2_WAVES
#define PSX_SYN2_WAVES 5
#endif
constexpr int kSyn2Pairs = PSX_SYN2_PAIRS;   // ring pairs per spin-2 synthesis thread
#ifndef PSX_AN2_PAIRS
#define PSX_AN2_PAIRS 4
#endif
#ifndef PSX_AN2_WAVES
#define PSX_AN2_WAVES 3
#endif
constexpr int kAn2Pairs = PSX_AN2_PAIRS;     // ring pairs per spin-2 analysis thread
constexpr int kSynChunk = 4;                // l values read ahead per synthesis round
constexpr double kBig = 0x1p600;            // rescaling threshold
constexpr double kInvBig = 0x1p-600;
constexpr double kLnBig = 415.88830833596718565;  // 600 ln 2
constexpr double kPi = 3.14159265358979323846264338327950288;

// ---- geometry -----------------------------------------------------------
struct Rings {
  int nside = 0, nr = 0, npairs = 0;
  int64_t npix = 0, xlen = 0;
  std::vector<int> n;
  std::vector<int64_t> first, xoff;
  std::vector<double> z, logsin, phi0;
};

Rings make_rings(int ns) {
  Rings R;
  R.nside = ns;
  R.nr = 4 * ns - 1;
  R.npairs = 2 * ns;
  R.npix = 12LL * ns * ns;
  R.n.resize(R.nr);
  R.first.resize(R.nr);
  R.xoff.resize(R.nr);
  R.z.resize(R.nr);
  R.logsin.resize(R.nr);
  R.phi0.resize(R.nr);
  const long double nsl = ns;
  int64_t off = 0;
  for (int k = 0; k < R.nr; ++k) {
    const int i = k + 1;
    long double z, phi0;
    if (i < ns) {
      R.n[k] = 4 * i;
      R.first[k] = 2LL * i * (i - 1);
      z = 1.0L - (long double)i * i / (3.0L * nsl * nsl);
      phi0 = (long double)kPi / (4.0L * i);
    } else if (i <= 3 * ns) {
      R.n[k] = 4 * ns;
      R.first[k] = 2LL * ns * (ns - 1) + (int64_t)(i - ns) * 4 * ns;
      z = 4.0L / 3.0L - 2.0L * i / (3.0L * nsl);
      phi0 = ((i - ns + 1) % 2) ? (long double)kPi / (4.0L * ns) : 0.0L;
    } else {
      const int ii = 4 * ns - i;
      R.n[k] = 4 * ii;
      R.first[k] = R.npix - 2LL * ii * (ii + 1);
      z = -(1.0L - (long double)ii * ii / (3.0L * nsl * nsl));
      phi0 = (long double)kPi / (4.0L * ii);
    }
    R.z[k] = (double)z;
    const long double s2 = (1.0L - z) * (1.0L + z);
    R.logsin[k] = (double)(0.5L * logl(s2));
    R.phi0[k] = (double)phi0;
    R.xoff[k] = off;
    off += R.n[k] / 2 + 1;
  }
  R.xlen = off;
  return R;
}

__host__ __device__ inline int64_t alm_idx(int lmax, int l, int m) {
  return (int64_t)m * (2 * lmax + 1 - m) / 2 + l;
}

// ---- kernels ------------------------------------------------------------
struct RingArgs {
  const int* n;
  const int64_t* first;
  const int64_t* xoff;
  const double* z;
  const double* logsin;
  const double* phi0;
  int nr, npairs, lmax;
};

// Forward DFT of one polar ring (n = 4i <= 4 nside - 4 pixels):
// X[q] = sum_j x_j e^{-2 pi i q j / n}, q = 0 .. n/2.  With the real input
// folded as s_j = x_j + x_{n-j}, d_j = x_j - x_{n-j} (cos is even, sin odd):
//   X[q] = x_0 + (-1)^q x_{n/2} + sum_{j=1}^{n/2-1} [s_j cos(2 pi q j/n) - i d_j sin(...)],
// half the terms of the direct sum.  Along j the (cos, sin) of a frequency
// advance by the angle-addition recurrence and are reset from the LDS cosine
// table every kRotAnchor steps, so rounding drift stays at ~kRotAnchor ulp.
constexpr int kRotAnchor = 32;
// Frequencies (or output pixel indices) per thread of the polar DFTs: one read
// of a step's four sub-sequence values from LDS feeds kPolarPer rotations.
constexpr int kPolarPer = 4;

// cos and sin of 2 pi t / n, n = 4i, 0 <= t < n, from the quarter-wave table
// cq[r] = cos(2 pi r / n), r = 0 .. i (sin(2 pi r / n) = cq[i - r]).
__device__ __forceinline__ void trig_q(const double* cq, int i, int t, double& c, double& s) {
  const bool h2 = t >= 2 * i;
  int r = h2 ? t - 2 * i : t;
  const bool h1 = r >= i;
  r = h1 ? r - i : r;
  const double a = cq[r], b = cq[i - r];
  const double c0 = h1 ? -b : a, s0 = h1 ? a : b;
  c = h2 ? -c0 : c0;
  s = h2 ? -s0 : s0;
}

// Polar ring of workgroup b: the largest rings first (ordinal ns - 2 - b/2,
// north and south alternating), so the longest workgroups start first.
__device__ __forceinline__ int polar_ring(const RingArgs& R, int ns, int b) {
  const int ord = ns - 2 - (b >> 1);
  return (b & 1) ? R.nr - 1 - ord : ord;
}

// Polar rings have n = 4i pixels, so the forward DFT splits in radix 4: the
// four real sub-sequences x_r[j] = x[4j + r] (j < i) have length-i DFTs
//   Y_r[p] = x_r[0] + (-1)^p x_r[i/2] (i even)
//            + sum_{j=1}^{ceil(i/2)-1} [s_rj cos(2 pi p j/i) - i d_rj sin(2 pi p j/i)]
// (s_rj, d_rj = x_r[j] +- x_r[i-j]); with Z_r = e^{-2 pi i r p/n} Y_r[p] and
// Y_r[i-p] = conj(Y_r[p]) (real input), one p gives four outputs:
//   X[p] = sum_r Z_r,                 X[i+p] = sum_r (-i)^r Z_r,
//   X[i-p] = sum_r (-i)^r conj(Z_r),  X[2i-p] = sum_r (-1)^r conj(Z_r),
// which cover X[0 .. 2i] for p = 0 .. floor(i/2) (each written once).  A
// thread takes kPolarPer consecutive p, each with its own (cos, sin)
// rotation; the sub-sequence values of a step are the same for every thread
// (LDS broadcast, two 16-byte reads each for s and -d).  The workgroup's two
// halves of NMAX/32 threads (the p blocks of the longest ring, i <= NMAX/4 - 1)
// sum the first and second half of the anchor blocks of j; the second half's
// sums then pass through LDS to the first, which writes the outputs.
template <int NMAX>
__global__ void __launch_bounds__(NMAX / 16) polar_dft(RingArgs R, const double* __restrict__ map,
                                                       double2* __restrict__ X, int ns) {
  constexpr int kHalf = NMAX / 32;
  static_assert(2 * (NMAX / 8) * 4 >= kHalf * 8 * kPolarPer, "sum exchange fits in sdv");
  __shared__ double cq[NMAX / 4 + 1];
  __shared__ __align__(16) double sdv[2][NMAX / 8][4];
  double (*const sv)[4] = sdv[0];                   // s_rj at [j][r], j >= 1
  double (*const dv)[4] = sdv[1];                   // -d_rj
  const int k = polar_ring(R, ns, blockIdx.x);
  const int n = R.n[k], i = n / 4;
  const int H = (i + 1) / 2 - 1;                    // folded pairs j = 1 .. H
  const double* x = map + R.first[k];
  for (int t = threadIdx.x; t <= i; t += blockDim.x) cq[t] = cospi(2.0 * t / n);
  for (int e = threadIdx.x; e < 4 * H; e += blockDim.x) {
    const int r = e & 3, j = (e >> 2) + 1;
    const double a = x[4 * j + r], b = x[4 * (i - j) + r];
    sv[j][r] = a + b;
    dv[j][r] = b - a;
  }
  __syncthreads();
  const bool even = (i & 1) == 0;
  double x0[4], xm[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    x0[r] = x[r];
    xm[r] = even ? x[4 * (i / 2) + r] : 0.0;
  }
  double2* out = X + R.xoff[k];
  const int P = i / 2;
  const int half = threadIdx.x / kHalf, lt = threadIdx.x % kHalf;
  const int pb = kPolarPer * lt;                    // p = pb .. pb + kPolarPer - 1
  const bool live = pb <= P;
  const int na = (H + kRotAnchor - 1) / kRotAnchor, na0 = (na + 1) / 2;
  const int jb = 1 + (half ? na0 : 0) * kRotAnchor;
  const int je = half ? H : min(H, na0 * kRotAnchor);
  double ar[kPolarPer][4], ai[kPolarPer][4];
  if (live) {
    double cd[kPolarPer], sd[kPolarPer], c[kPolarPer], sn[kPolarPer];
#pragma unroll
    for (int u = 0; u < kPolarPer; ++u) {
      const int p = pb + u;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        ar[u][r] = half ? 0.0 : x0[r] + ((p & 1) ? -xm[r] : xm[r]);
        ai[u][r] = 0.0;
      }
      trig_q(cq, i, (4 * p) % n, cd[u], sd[u]);     // one step: angle 2 pi p / i
    }
    for (int j0 = jb; j0 <= je; j0 += kRotAnchor) {
#pragma unroll
      for (int u = 0; u < kPolarPer; ++u) trig_q(cq, i, (4 * (pb + u) * j0) % n, c[u], sn[u]);
      const int jend = min(je + 1, j0 + kRotAnchor);
      for (int j = j0; j < jend; ++j) {
        const double2 s01 = *(const double2*)&sv[j][0], s23 = *(const double2*)&sv[j][2];
        const double2 d01 = *(const double2*)&dv[j][0], d23 = *(const double2*)&dv[j][2];
        const double sj[4] = {s01.x, s01.y, s23.x, s23.y};
        const double dj[4] = {d01.x, d01.y, d23.x, d23.y};
#pragma unroll
        for (int u = 0; u < kPolarPer; ++u) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            ar[u][r] = fma(sj[r], c[u], ar[u][r]);
            ai[u][r] = fma(dj[r], sn[u], ai[u][r]);
          }
          const double c2 = fma(c[u], cd[u], -sn[u] * sd[u]);
          const double s2 = fma(sn[u], cd[u], c[u] * sd[u]);
          c[u] = c2;
          sn[u] = s2;
        }
      }
    }
  }
  // second half -> first half through sdv (no longer read)
  __syncthreads();
  double* xs = &sdv[0][0][0];
  if (half && live) {
#pragma unroll
    for (int u = 0; u < kPolarPer; ++u)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        xs[((2 * u) * 4 + r) * kHalf + lt] = ar[u][r];
        xs[((2 * u + 1) * 4 + r) * kHalf + lt] = ai[u][r];
      }
  }
  __syncthreads();
  if (half || !live) return;
#pragma unroll
  for (int u = 0; u < kPolarPer; ++u)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      ar[u][r] += xs[((2 * u) * 4 + r) * kHalf + lt];
      ai[u][r] += xs[((2 * u + 1) * 4 + r) * kHalf + lt];
    }
  {
#pragma unroll
    for (int u = 0; u < kPolarPer; ++u) {
      const int p = pb + u;
      if (p > P) break;
      // Z_r = e^{-2 pi i r p / n} Y_r  (r p < n)
      double zx[4], zy[4];
      zx[0] = ar[u][0];
      zy[0] = ai[u][0];
#pragma unroll
      for (int r = 1; r < 4; ++r) {
        double cr, sr;
        trig_q(cq, i, r * p, cr, sr);
        zx[r] = fma(cr, ar[u][r], sr * ai[u][r]);
        zy[r] = fma(cr, ai[u][r], -sr * ar[u][r]);
      }
      const double2 Xp = make_double2(zx[0] + zx[1] + zx[2] + zx[3], zy[0] + zy[1] + zy[2] + zy[3]);
      const double2 Xip = make_double2(zx[0] + zy[1] - zx[2] - zy[3], zy[0] - zx[1] - zy[2] + zx[3]);
      const double2 Xim = make_double2(zx[0] - zy[1] - zx[2] + zy[3], -zy[0] - zx[1] + zy[2] + zx[3]);
      const double2 X2m = make_double2(zx[0] - zx[1] + zx[2] - zx[3], -zy[0] + zy[1] - zy[2] + zy[3]);
      if (p == 0) {
        out[0] = Xp;
        out[i] = Xip;
        out[2 * i] = X2m;
      } else if (2 * p < i) {
        out[p] = Xp;
        out[i - p] = Xim;
        out[i + p] = Xip;
        out[2 * i - p] = X2m;
      } else {                                      // p = i / 2
        out[p] = Xp;
        out[i + p] = Xip;
      }
    }
  }
}

// Inverse DFT of one polar ring (n = 4i pixels) from its Hermitian half
// Y[0 .. n/2] (Y_0, Y_{n/2} taken as real): f_j = sum_{q<n} Yh_q e^{2 pi i q j/n}
// with Yh_q = Y_q (q <= n/2), conj(Y_{n-q}) above.  Radix 4 on the output,
// j = 4m + r:
//   W_r[p] = e^{2 pi i p r/n} sum_{k<4} i^{kr} Yh_{p+ki},
//   f_{4m+r} = sum_{p<i} W_r[p] e^{2 pi i p m/i}
//            = W_r[0] + (-1)^m W_r[i/2] (i even)
//              + 2 sum_{p=1}^{ceil(i/2)-1} Re(W_r[p] e^{2 pi i p m/i})
// (W_r is Hermitian because f is real).  A thread takes kPolarPer
// consecutive m, each with its own rotation, and writes pixels 4m + r and
// 4(i - m) + r, whose cosine sums agree and sine sums are opposite.  As in
// the forward DFT, the workgroup's two halves sum the two halves of the
// anchor blocks of p and meet in LDS.
template <int NMAX>
__global__ void __launch_bounds__(NMAX / 16) polar_idft(RingArgs R, const double2* __restrict__ X,
                                                        double* __restrict__ map, int ns) {
  constexpr int kHalf = NMAX / 32;
  static_assert(2 * (NMAX / 8 + 1) * 4 >= kHalf * 8 * kPolarPer, "sum exchange fits in wxy");
  __shared__ double cq[NMAX / 4 + 1];
  __shared__ __align__(16) double wxy[2][NMAX / 8 + 1][4];
  double (*const wx)[4] = wxy[0];                   // Re 2 W_r[p] at [p][r]
  double (*const wy)[4] = wxy[1];                   // Im 2 W_r[p]
  const int k = polar_ring(R, ns, blockIdx.x);
  const int n = R.n[k], h = n / 2, i = n / 4;
  for (int t = threadIdx.x; t <= i; t += blockDim.x) cq[t] = cospi(2.0 * t / n);
  __syncthreads();
  const double2* Y = X + R.xoff[k];
  const int P = i / 2;
  for (int p = threadIdx.x; p <= P; p += blockDim.x) {
    double2 y[4];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      const int q = p + kk * i;
      if (q <= h) {
        y[kk] = Y[q];
        if (q == 0 || q == h) y[kk].y = 0.0;
      } else {
        y[kk] = Y[n - q];
        y[kk].y = -y[kk].y;
      }
    }
    // sum_k i^{kr} y_k for r = 0 .. 3 (i (a + ib) = -b + ia)
    double bx[4], by[4];
    bx[0] = y[0].x + y[1].x + y[2].x + y[3].x;
    by[0] = y[0].y + y[1].y + y[2].y + y[3].y;
    bx[1] = y[0].x - y[1].y - y[2].x + y[3].y;
    by[1] = y[0].y + y[1].x - y[2].y - y[3].x;
    bx[2] = y[0].x - y[1].x + y[2].x - y[3].x;
    by[2] = y[0].y - y[1].y + y[2].y - y[3].y;
    bx[3] = y[0].x + y[1].y - y[2].x - y[3].y;
    by[3] = y[0].y - y[1].x - y[2].y + y[3].x;
    wx[p][0] = 2.0 * bx[0];
    wy[p][0] = 2.0 * by[0];
#pragma unroll
    for (int r = 1; r < 4; ++r) {
      double cr, sr;                                // e^{+2 pi i p r / n}  (r p < n)
      trig_q(cq, i, r * p, cr, sr);
      wx[p][r] = 2.0 * fma(cr, bx[r], -sr * by[r]);
      wy[p][r] = 2.0 * fma(cr, by[r], sr * bx[r]);
    }
  }
  __syncthreads();
  double* f = map + R.first[k];
  const bool even = (i & 1) == 0;
  const int H = (i + 1) / 2 - 1;                    // p = 1 .. H
  double w0[4], wP[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    w0[r] = wx[0][r];
    wP[r] = even ? wx[P][r] : 0.0;
  }
  const int half = threadIdx.x / kHalf, lt = threadIdx.x % kHalf;
  const int mb = kPolarPer * lt;                    // m = mb .. mb + kPolarPer - 1
  const bool live = mb <= P;
  const int na = (H + kRotAnchor - 1) / kRotAnchor, na0 = (na + 1) / 2;
  const int qb = 1 + (half ? na0 : 0) * kRotAnchor;
  const int qe = half ? H : min(H, na0 * kRotAnchor);
  double A[kPolarPer][4], B[kPolarPer][4];
  if (live) {
    double cd[kPolarPer], sd[kPolarPer], c[kPolarPer], sn[kPolarPer];
#pragma unroll
    for (int u = 0; u < kPolarPer; ++u) {
      const int m = mb + u;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        A[u][r] = half ? 0.0 : 0.5 * (w0[r] + ((m & 1) ? -wP[r] : wP[r]));
        B[u][r] = 0.0;
      }
      trig_q(cq, i, (4 * m) % n, cd[u], sd[u]);     // one step: angle 2 pi m / i
    }
    for (int p0 = qb; p0 <= qe; p0 += kRotAnchor) {
#pragma unroll
      for (int u = 0; u < kPolarPer; ++u) trig_q(cq, i, (4 * (mb + u) * p0) % n, c[u], sn[u]);
      const int pend = min(qe + 1, p0 + kRotAnchor);
      for (int p = p0; p < pend; ++p) {
        const double2 x01 = *(const double2*)&wx[p][0], x23 = *(const double2*)&wx[p][2];
        const double2 y01 = *(const double2*)&wy[p][0], y23 = *(const double2*)&wy[p][2];
        const double px[4] = {x01.x, x01.y, x23.x, x23.y};
        const double py[4] = {y01.x, y01.y, y23.x, y23.y};
#pragma unroll
        for (int u = 0; u < kPolarPer; ++u) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            A[u][r] = fma(px[r], c[u], A[u][r]);
            B[u][r] = fma(py[r], sn[u], B[u][r]);
          }
          const double c2 = fma(c[u], cd[u], -sn[u] * sd[u]);
          const double s2 = fma(sn[u], cd[u], c[u] * sd[u]);
          c[u] = c2;
          sn[u] = s2;
        }
      }
    }
  }
  // second half -> first half through wxy (no longer read)
  __syncthreads();
  double* xs = &wxy[0][0][0];
  if (half && live) {
#pragma unroll
    for (int u = 0; u < kPolarPer; ++u)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        xs[((2 * u) * 4 + r) * kHalf + lt] = A[u][r];
        xs[((2 * u + 1) * 4 + r) * kHalf + lt] = B[u][r];
      }
  }
  __syncthreads();
  if (half || !live) return;
#pragma unroll
  for (int u = 0; u < kPolarPer; ++u) {
    const int m = mb + u;
    if (m > P) break;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const double a = A[u][r] + xs[((2 * u) * 4 + r) * kHalf + lt];
      const double b = B[u][r] + xs[((2 * u + 1) * 4 + r) * kHalf + lt];
      f[4 * m + r] = a - b;
      if (m > 0 && 2 * m < i) f[4 * (i - m) + r] = a + b;
    }
  }
}

// F+-[m][p] from the ring spectra (north ring p, south ring nr-1-p; the
// equator pairs with itself and counts once).  The two rings of a pair have
// the same n and phi0 (pi/n or 0), so with m = a n + q (0 <= q < n) the
// shift e^{-i m phi0} = s^a e^{-i q phi0} (s = -1 where phi0 != 0) is one
// small-angle sincos for both; X[q] = conj(X[n - q]) above n/2.
// A workgroup takes kFaPairs ring pairs and kFaM consecutive m: each wave
// reads four pairs' spectra along m (consecutive q: one cache line serves
// eight lanes), the sums pass through LDS, and F+- leave as runs of
// kFaPairs pairs per m.
constexpr int kFaPairs = 16, kFaM = 64;

__global__ void __launch_bounds__(256) fold_analysis(RingArgs R, const double2* __restrict__ X,
                                                     double2* __restrict__ Fp,
                                                     double2* __restrict__ Fm) {
  __shared__ double2 tp[kFaM][kFaPairs + 1], tm[kFaM][kFaPairs + 1];
  const int p0 = (int)blockIdx.x * kFaPairs, m0 = (int)blockIdx.y * kFaM;
  const int lane = (int)threadIdx.x & 63, wv = (int)threadIdx.x >> 6;
  const int m = m0 + lane;
#pragma unroll
  for (int k = 0; k < kFaPairs / 4; ++k) {
    const int pl = wv * (kFaPairs / 4) + k, p = p0 + pl;
    double2 A = make_double2(0.0, 0.0), B = make_double2(0.0, 0.0);
    if (p < R.npairs && m <= R.lmax) {
      const int kn = p, ks = R.nr - 1 - p;
      const int n = R.n[kn];
      const int a = m / n, q = m - a * n;
      const double phi0 = R.phi0[kn];
      double sn, c;
      sincos((double)q * phi0, &sn, &c);
      if (phi0 != 0.0 && (a & 1)) {
        sn = -sn;
        c = -c;
      }
      const bool hi = 2 * q > n;
      const int qi = hi ? n - q : q;
      const double cy = hi ? -1.0 : 1.0;
      // v e^{-i m phi0}
      auto coef = [&](int k2) {
        const double2 v = X[R.xoff[k2] + qi];
        const double vy = cy * v.y;
        return make_double2(v.x * c + vy * sn, vy * c - v.x * sn);
      };
      A = coef(kn);
      if (ks != kn) B = coef(ks);
    }
    tp[lane][pl] = make_double2(A.x + B.x, A.y + B.y);
    tm[lane][pl] = make_double2(A.x - B.x, A.y - B.y);
  }
  __syncthreads();
  const int pl = (int)threadIdx.x % kFaPairs, p = p0 + pl;
#pragma unroll
  for (int jj = 0; jj < kFaM * kFaPairs / 256; ++jj) {
    const int ml = (int)threadIdx.x / kFaPairs + (256 / kFaPairs) * jj, mm = m0 + ml;
    if (p < R.npairs && mm <= R.lmax) {
      const int64_t o = (int64_t)mm * R.npairs + p;
      Fp[o] = tp[ml][pl];
      Fm[o] = tm[ml][pl];
    }
  }
}

// Recurrence coefficients per (l, m), l >= m + 2, in a_lm order:
//   lambda_l = A (z lambda_{l-1} - B lambda_{l-2}),
//   A = sqrt((4l^2 - 1)/(l^2 - m^2)),  B = sqrt(((l-1)^2 - m^2)/(4(l-1)^2 - 1)).
__global__ void __launch_bounds__(256) rec_coefs(int lmax, double2* __restrict__ AB) {
  const int m = blockIdx.y;
  const int l = m + blockIdx.x * blockDim.x + threadIdx.x;
  if (l > lmax) return;
  double a = 0.0, b = 0.0;
  if (l >= m + 2) {
    const double L = l, M = m;
    a = sqrt((4.0 * L * L - 1.0) / ((L - M) * (L + M)));
    b = sqrt(((L - 1.0 - M) * (L - 1.0 + M)) / (4.0 * (L - 1.0) * (L - 1.0) - 1.0));
  } else if (l == m + 1) {
    a = sqrt(2.0 * m + 3.0);
  }
  AB[alm_idx(lmax, l, m)] = make_double2(a, b);
}

struct LegArgs {
  RingArgs R;
  const double2* AB;      // recurrence coefficients
  const double* logc;     // log of sqrt((2m+1)/(4 pi) prod (2k-1)/(2k)), per m
  int nchunk;             // pair chunks
};

// Scaled lambda_mm for ring k: mantissa and scale s <= 0 (value = mant 2^(600 s)).
__device__ inline void lambda_start(const LegArgs& A, int k, int m, double& mant, int& s) {
  const double L = A.logc[m] + (double)m * A.R.logsin[k];
  const int sc = L >= 0.0 ? 0 : -(int)floor(-L / kLnBig);
  mant = exp(L - (double)sc * kLnBig);
  if (m & 1) mant = -mant;
  s = sc;
}

// Uniform load through the constant address space: the value lands in SGPRs
// (every lane reads the same address).
__device__ __forceinline__ double sload(const double* p) {
  return *((const __attribute__((address_space(4))) double*)p);
}
__device__ __forceinline__ double2 sload(const double2* p) {
  const double* q = reinterpret_cast<const double*>(p);
  return make_double2(sload(q), sload(q + 1));
}

// Weight of a scaled value mant 2^(600 s) in a sum: 1, 2^-600, or 0 below
// 2^-1200 (too small to matter next to O(1) terms).
__device__ __forceinline__ double scale_weight(int s) {
  return s == 0 ? 1.0 : (s == -1 ? kInvBig : 0.0);
}

// Rescaling runs once per chunk of kLChunk steps, so the steps themselves
// carry no branch: from a mantissa below 2^600 the recurrence grows by less
// than 2^48 over 8 steps (sqrt(2m+3) sqrt(m) sqrt(2m/3) ... < 2^48 for
// m <= 8191), far from double overflow, and a value that crosses 2^600 inside
// a chunk is still weighted exactly (or dropped when it is below 2^-600).
__device__ __forceinline__ void rescale_pair(double& a, double& b, int& s, double& wgt) {
  if (s < 0 && fabs(a) > kBig) {
    a *= kInvBig;
    b *= kInvBig;
    s += 1;
    wgt = scale_weight(s);
  }
}

// Up to kMaxGroup maps share one pass of the recurrences: lambda_lm(z) is the
// same for every map, only the accumulations are per map.
constexpr int kMaxGroup = 4;

template <int NM>
struct MapPtrs {
  const double2* in[NM];   // analysis: F+ of each map; synthesis: a_lm
  const double2* in2[NM];  // analysis: F- of each map
  double2* out[NM];        // analysis: partial sums; synthesis: G
};

// l values per LDS reduction round of the analysis: 2 NM values per l and
// thread go through the round's 16 rows of 256 doubles (32 KB)
__host__ __device__ constexpr int an_lchunk(int nm) { return nm == 1 ? 8 : nm == 2 ? 4 : 2; }

// DPP move of a double inside a row of 16 lanes (both 32-bit halves).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
  const long long v = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(v & 0xffffffffLL), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(v >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the 16 lanes of a row: quad_perm [1,0,3,2], quad_perm [2,3,0,1],
// row_ror 4, row_ror 8.  Every lane ends with the row's sum (lanes of quads
// 0, 2 and of quads 1, 3 may differ in the last bit: the caller takes lane 0).
__device__ __forceinline__ double row16_sum(double s) {
  s += dpp_f64<0xB1>(s);
  s += dpp_f64<0x4E>(s);
  s += dpp_f64<0x124>(s);
  s += dpp_f64<0x128>(s);
  return s;
}

// Analysis partial sums: part_q[chunk][idx(l, m)] = sum over the chunk's pairs
// of lambda_lm(z_p) F+-_p(m) for each map q of the group.  Each thread carries
// kPairsPerThread independent recurrences; per round of LC l values the 256
// partial sums of every (l, map, re/im) are reduced through LDS (16 lanes per
// output, each summing a stride-16 column of its row so that a wave's reads
// are consecutive doubles, then a DPP sum over the 16 lanes), in the same
// order whatever the group size (one map alone gives the same sums).  The
// round holding l = m, m + 1 and the last, partial round take the checked
// path; the rounds between carry no per-step branch.
template <int NM>
__global__ void __launch_bounds__(256) legendre_analysis(LegArgs A, MapPtrs<NM> M) {
  constexpr int LC = an_lchunk(NM);
  constexpr int NO = 2 * LC * NM;  // outputs per round
  static_assert(NO <= 16, "one output per row of 16 lanes");
  constexpr int P = kPairsPerThread;
  __shared__ double red[NO][256];
  const int m = blockIdx.y;
  const int chunk = blockIdx.x;
  const int lmax = A.R.lmax;
  const int tid = threadIdx.x;
  const double2* ABm = A.AB + alm_idx(lmax, m, m);   // ABm[l - m]
  double z[P], l1[P], l2[P], wgt[P];
  int sc[P];
  double2 fp[NM][P], fm[NM][P];
#pragma unroll
  for (int r = 0; r < P; ++r) {
    const int p = (chunk * P + r) * 256 + tid;
    const bool live = p < A.R.npairs;
    const int pp = live ? p : 0;
    z[r] = A.R.z[pp];
    lambda_start(A, pp, m, l1[r], sc[r]);
    wgt[r] = scale_weight(sc[r]);
    l2[r] = 0.0;
    const int64_t o = (int64_t)m * A.R.npairs + pp;
#pragma unroll
    for (int q = 0; q < NM; ++q) {
      fp[q][r] = live ? M.in[q][o] : make_double2(0.0, 0.0);
      fm[q][r] = live ? M.in2[q][o] : make_double2(0.0, 0.0);
    }
  }
  const int o = tid >> 4, g = tid & 15;
  const int oj = o / (2 * NM), oq = (o >> 1) % NM, oc = o & 1;
  const int64_t nalm = alm_idx(lmax, lmax + 1, lmax + 1);
  double2* outq = M.out[0];
#pragma unroll
  for (int q = 1; q < NM; ++q)
    if (oq == q) outq = M.out[q];
  outq += (int64_t)chunk * nalm;
  auto round = [&](const int l0, const bool edge) {
#pragma unroll
    for (int j = 0; j < LC; ++j) {
      const int l = l0 + j;
      double accr[NM], acci[NM];
#pragma unroll
      for (int q = 0; q < NM; ++q) accr[q] = acci[q] = 0.0;
      if (!edge || l <= lmax) {
        const double2 ab = sload(ABm + (l - m));   // (0, 0) at l = m
        const bool odd = j & 1;                    // l0 - m is a multiple of LC
#pragma unroll
        for (int r = 0; r < P; ++r) {
          double lam;
          if (edge && l == m) {
            lam = l1[r];
          } else {
            lam = (edge && l == m + 1) ? ab.x * z[r] * l1[r]
                                       : ab.x * fma(z[r], l1[r], -ab.y * l2[r]);
            l2[r] = l1[r];
            l1[r] = lam;
          }
          const double lt = lam * wgt[r];
#pragma unroll
          for (int q = 0; q < NM; ++q) {
            const double2 f = odd ? fm[q][r] : fp[q][r];
            accr[q] = fma(lt, f.x, accr[q]);
            acci[q] = fma(lt, f.y, acci[q]);
          }
        }
      }
#pragma unroll
      for (int q = 0; q < NM; ++q) {
        red[(j * NM + q) * 2][tid] = accr[q];
        red[(j * NM + q) * 2 + 1][tid] = acci[q];
      }
    }
#pragma unroll
    for (int r = 0; r < P; ++r) rescale_pair(l1[r], l2[r], sc[r], wgt[r]);
    __syncthreads();
    if (o < NO) {                  // whole rows of 16 lanes
      double h[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double* row = &red[o][64 * k + g];
        h[k] = (row[0] + row[16]) + (row[32] + row[48]);
      }
      const double s = row16_sum((h[0] + h[1]) + (h[2] + h[3]));
      const int l = l0 + oj;
      if (g == 0 && l <= lmax) ((double*)(outq + alm_idx(lmax, l, m)))[oc] = s;
    }
    __syncthreads();
  };
  int l0 = m;
  round(l0, true);
  for (l0 += LC; l0 + LC - 1 <= lmax; l0 += LC) round(l0, false);
  if (l0 <= lmax) round(l0, true);
}

// alm (+)= w sum_chunk part[chunk]
__global__ void __launch_bounds__(256) sum_partials(const double2* __restrict__ part,
                                                    int nchunk, int64_t nalm, double w,
                                                    int accumulate, double2* __restrict__ alm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nalm) return;
  double sr = 0.0, si = 0.0;
  for (int c = 0; c < nchunk; ++c) {
    const double2 v = part[(int64_t)c * nalm + i];
    sr += v.x;
    si += v.y;
  }
  double2 a = accumulate ? alm[i] : make_double2(0.0, 0.0);
  a.x = fma(w, sr, a.x);
  a.y = fma(w, si, a.y);
  alm[i] = a;
}

// G, the Legendre synthesis's output (G[m][ring] = sum_l a_lm lambda_lm), is
// stored in groups of kGRings rings: entry (m, k) at
// ((k / kGRings) (lmax + 1) + m) kGRings + k % kGRings.  The syntheses write
// 128-byte runs of 8 adjacent rings, and the fold, which takes 8 adjacent
// rings per workgroup, reads whole rows of a group for consecutive m.
constexpr int kGRings = 8;
__host__ __device__ inline int64_t gidx(int lmax, int m, int k) {
  return ((int64_t)(k / kGRings) * (lmax + 1) + m) * kGRings + (k % kGRings);
}
// entries of G: the ring count padded to whole groups
inline int64_t g_elems(int lmax, int nr) {
  return (int64_t)((nr + kGRings - 1) / kGRings) * kGRings * (lmax + 1);
}

// ring pairs per synthesis thread: fewer for larger groups (per pair and
// map, four accumulators)
#ifndef PSX_SYN_PAIRS4
#define PSX_SYN_PAIRS4 2
#endif
#ifndef PSX_SYN_WAVES
#define PSX_SYN_WAVES 1
#endif
__host__ __device__ constexpr int syn_pairs(int nm) { return nm <= 2 ? kSynPairs : PSX_SYN_PAIRS4; }

// Synthesis: G_q[m][k] = sum_l a_q,lm lambda_lm(z_k) for SP ring pairs per
// thread and each map q of the group (north: even + odd l+m parts, south:
// even - odd); a_lm and the recurrence coefficients are uniform across the
// workgroup and read with scalar loads, kSynChunk at a time ahead of the
// steps that use them (a rescale check per such round: growth over 4 steps
// stays below 2^26).
template <int NM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PSX_SYN_WAVES)))
legendre_synthesis(LegArgs A, MapPtrs<NM> M) {
  constexpr int SP = syn_pairs(NM);
  const int m = blockIdx.y;
  const int lmax = A.R.lmax;
  const int64_t i0 = alm_idx(lmax, m, m);
  const double2* ABm = A.AB + i0;
  double z[SP], l1[SP], l2[SP], wgt[SP];
  double er[NM][SP], ei[NM][SP], orr[NM][SP], oi[NM][SP];
  int sc[SP], pr[SP];
#pragma unroll
  for (int r = 0; r < SP; ++r) {
    const int p = (blockIdx.x * SP + r) * blockDim.x + threadIdx.x;
    pr[r] = p;
    const int pp = p < A.R.npairs ? p : 0;
    z[r] = A.R.z[pp];
    lambda_start(A, pp, m, l1[r], sc[r]);
    wgt[r] = scale_weight(sc[r]);
    l2[r] = 0.0;
#pragma unroll
    for (int q = 0; q < NM; ++q) er[q][r] = ei[q][r] = orr[q][r] = oi[q][r] = 0.0;
  }
  // the first round holds l = m, m + 1, the last may pass lmax: both take
  // the checked path, the rounds between load and step unconditionally
  auto round = [&](const int l0, const bool edge) {
    double2 a[NM][kSynChunk], ab[kSynChunk];
#pragma unroll
    for (int j = 0; j < kSynChunk; ++j) {
      const bool in = !edge || l0 + j <= lmax;
#pragma unroll
      for (int q = 0; q < NM; ++q)
        a[q][j] = in ? sload(M.in[q] + i0 + (l0 + j - m)) : make_double2(0.0, 0.0);
      ab[j] = in ? sload(ABm + (l0 + j - m)) : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int j = 0; j < kSynChunk; ++j) {
      const int l = l0 + j;
      if (edge && l > lmax) break;
      const bool odd = j & 1;                      // l0 - m is a multiple of kSynChunk
#pragma unroll
      for (int r = 0; r < SP; ++r) {
        double lam;
        if (edge && l == m) {
          lam = l1[r];
        } else {
          lam = (edge && l == m + 1) ? ab[j].x * z[r] * l1[r]
                                     : ab[j].x * fma(z[r], l1[r], -ab[j].y * l2[r]);
          l2[r] = l1[r];
          l1[r] = lam;
        }
        const double lt = lam * wgt[r];
#pragma unroll
        for (int q = 0; q < NM; ++q) {
          if (odd) {
            orr[q][r] = fma(lt, a[q][j].x, orr[q][r]);
            oi[q][r] = fma(lt, a[q][j].y, oi[q][r]);
          } else {
            er[q][r] = fma(lt, a[q][j].x, er[q][r]);
            ei[q][r] = fma(lt, a[q][j].y, ei[q][r]);
          }
        }
      }
    }
#pragma unroll
    for (int r = 0; r < SP; ++r) rescale_pair(l1[r], l2[r], sc[r], wgt[r]);
  };
  int l0 = m;
  round(l0, true);
  for (l0 += kSynChunk; l0 + kSynChunk - 1 <= lmax; l0 += kSynChunk) round(l0, false);
  if (l0 <= lmax) round(l0, true);
#pragma unroll
  for (int r = 0; r < SP; ++r) {
    const int p = pr[r];
    if (p >= A.R.npairs) continue;
    const int kn = p, ks = A.R.nr - 1 - p;
#pragma unroll
    for (int q = 0; q < NM; ++q) {
      double2* G = M.out[q];
      G[gidx(lmax, m, kn)] = make_double2(er[q][r] + orr[q][r], ei[q][r] + oi[q][r]);
      if (ks != kn)
        G[gidx(lmax, m, ks)] = make_double2(er[q][r] - orr[q][r], ei[q][r] - oi[q][r]);
    }
  }
}

// Hermitian half spectrum of each ring from G (layout of gidx).
// Y[q] collects G_m e^{i m phi0} for m = q mod n and conj(G_m e^{i m phi0})
// for m = -q mod n (m >= 1); m = 0 contributes Re(G_0).  Every HEALPix ring
// has phi0 = pi/n (first pixel shifted by half a pixel) or phi0 = 0, so
// e^{i n phi0} = s = -1 or +1 exactly and
//   Y[q] = e^{i q phi0} (A + conj(B)),  A = sum_{j>=0} s^j G_{q+jn},
//                                      B = sum_{j>=1} s^j G_{jn-q}:
// one small-angle sincos per output.  A workgroup covers kFoldRings adjacent
// rings (lane & 7) and kFoldQ values of q; where the rings share n, a wave
// reads 8 consecutive rows of one ring group (1 KB contiguous).  Workgroup b
// runs on XCD b mod 8: the 1-D grid gives all q-blocks of one ring group the
// same XCD, so the rows of G the group reads are fetched into one L2.
//
// A polar ring of n pixels folds about 2 (lmax + 1) / n terms into each
// output: ~1500 dependent loads for n = 4 at lmax 3071, which would make the
// few smallest rings the critical path of the whole launch.  The rings with
// n <= kFoldSmallN (the first nsmall rings of each cap) therefore get a
// workgroup each, launched ahead of the ring groups: each output is summed by a
// team of S lanes (S a power of two, S (n/2 + 1) <= 256, S <= 64), lane t of
// the team taking j = t, t + S, ..., and the team's partial sums are combined
// by a butterfly of fixed order.
constexpr int kFoldRings = kGRings, kFoldQ = 32, kFoldSmallN = 508;

__device__ __forceinline__ void fold_small_ring(const RingArgs& R, const double2* __restrict__ G,
                                                double2* __restrict__ X, int k) {
  const int n = R.n[k], nq = n / 2 + 1;
  int S = 64;
  while (S * nq > 256) S >>= 1;
  const int q = threadIdx.x / S, t = threadIdx.x % S;
  if (q >= nq) return;                              // whole teams: S divides 64
  const double phi0 = R.phi0[k];
  const bool neg = phi0 != 0.0;                     // s = -1
  double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
  for (int j = t, m = q + t * n; m <= R.lmax; j += S, m += S * n) {
    const double2 g = G[gidx(R.lmax, m, k)];
    const double sg = (neg && (j & 1)) ? -1.0 : 1.0;
    ar = fma(sg, g.x, ar);
    if (m != 0) ai = fma(sg, g.y, ai);
  }
  for (int j = 1 + t, m = (1 + t) * n - q; m <= R.lmax; j += S, m += S * n) {
    const double2 g = G[gidx(R.lmax, m, k)];
    const double sg = (neg && (j & 1)) ? -1.0 : 1.0;
    br = fma(sg, g.x, br);
    bi = fma(sg, g.y, bi);
  }
  for (int w = 1; w < S; w <<= 1) {
    ar += __shfl_xor(ar, w);
    ai += __shfl_xor(ai, w);
    br += __shfl_xor(br, w);
    bi += __shfl_xor(bi, w);
  }
  if (t != 0) return;
  double sn, c;
  sincos((double)q * phi0, &sn, &c);
  const double tr = ar + br, ti = ai - bi;
  X[R.xoff[k] + q] = make_double2(c * tr - sn * ti, sn * tr + c * ti);
}

__global__ void __launch_bounds__(256) fold_synthesis(RingArgs R, int nqb, int nsmall, int nbs,
                                                      const double2* __restrict__ G,
                                                      double2* __restrict__ X) {
  if ((int)blockIdx.x < nbs) {                      // one small polar ring per workgroup
    const int r = blockIdx.x;
    if (r < 2 * nsmall) fold_small_ring(R, G, X, r < nsmall ? r : R.nr - 1 - (r - nsmall));
    return;
  }
  const int b = blockIdx.x - nbs;                   // nbs is a multiple of 8: XCD b mod 8
  const int u = b >> 3;
  const int grp = (b & 7) + 8 * (u / nqb);
  const int k = grp * kFoldRings + (threadIdx.x & (kFoldRings - 1));
  const int q = (u % nqb) * kFoldQ + (threadIdx.x / kFoldRings);
  if (k >= R.nr || k < nsmall || k >= R.nr - nsmall) return;
  const int n = R.n[k];
  if (q > n / 2) return;
  const double phi0 = R.phi0[k];
  const double s = phi0 != 0.0 ? -1.0 : 1.0;
  double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0, sg = 1.0;
  int m = q;
  if (q == 0) {
    ar = G[gidx(R.lmax, 0, k)].x;
    m = n;
    sg = s;
  }
  for (; m <= R.lmax; m += n, sg *= s) {
    const double2 g = G[gidx(R.lmax, m, k)];
    ar = fma(sg, g.x, ar);
    ai = fma(sg, g.y, ai);
  }
  sg = s;
  for (m = n - q; m <= R.lmax; m += n, sg *= s) {
    const double2 g = G[gidx(R.lmax, m, k)];
    br = fma(sg, g.x, br);
    bi = fma(sg, g.y, bi);
  }
  double sn, c;
  sincos((double)q * phi0, &sn, &c);
  const double tr = ar + br, ti = ai - bi;
  X[R.xoff[k] + q] = make_double2(c * tr - sn * ti, sn * tr + c * ti);
}

void launch_fold_synthesis(const Rings& R, hipStream_t s, RingArgs ra, const double2* G,
                           double2* X) {
  const int nqb = (2 * R.nside + 1 + kFoldQ - 1) / kFoldQ;     // n/2 <= 2 nside
  const int ngrp = (R.nr + kFoldRings - 1) / kFoldRings;
  // polar rings k < nside - 1 have n = 4 (k + 1) pixels
  const int nsmall = std::min(R.nside - 1, kFoldSmallN / 4);
  const int nbs = (2 * nsmall + 7) / 8 * 8;
  const int64_t nb = nbs + (int64_t)8 * nqb * ((ngrp + 7) / 8);
  hipLaunchKernelGGL(fold_synthesis, dim3((unsigned)nb), dim3(256), 0, s, ra, nqb, nsmall, nbs,
                     G, X);
  PSX_HIP(hipGetLastError());
}

// ---- spin 2 ---------------------------------------------------------------
// With _sY_lm = (-1)^s N_l d^l_{m,-s}(theta) e^{i m phi}, N_l = sqrt((2l+1)/(4 pi)),
// the two real functions per (l, m, ring) are X = N_l d^l_{m,-2} and
// Y = N_l d^l_{m,2}; lambda+- = (X +- Y)/2.  For Fourier coefficients Q_m, U_m
// of the real Q and U rings (HEALPix: Q +- iU = sum _{+-2}a _{+-2}Y,
// a_E = -(_2a + _-2a)/2, a_B = i(_2a - _-2a)/2):
//   a_E = -w sum_rings [lambda+ Q_m + i lambda- U_m],
//   a_B = -w sum_rings [lambda+ U_m - i lambda- Q_m],
// and synthesis is the transpose,
//   Q_m = -sum_l [lambda+ a_E + i lambda- a_B],  U_m = -sum_l [lambda+ a_B - i lambda- a_E].
// d^l_{m,m'}(-z) = (-1)^{l+m} d^l_{m,-m'}(z): lambda+ has the ring-pair
// parity (-1)^{l+m} of the spin-0 case, lambda- the opposite one, so the
// spin-0 fold (F+ = F_N + F_S, F- = F_N - F_S) serves Q and U unchanged.
// X and Y each follow a three-term recurrence in l (they differ in the sign
// of the m m' = -+2m term); run separately, each keeps full relative
// precision, while a recurrence on lambda+- directly loses the small one near
// the pole.
constexpr int kLChunk2 = 4;                 // l values per LDS round, spin 2

struct PolArgs {
  RingArgs R;
  const double* omz;      // 1 - z per ring, exact near the north pole
  const double* abc;      // recurrence coefficients, 3 per (l, m)
  const double* logc2;    // log(N_m K_m) - (m - 2) ln 2 per m (m >= 2)
  int nchunk;
};

// Per (l, m), in a_lm order, the coefficients producing degree l from l - 1
// and l - 2 (L = l - 1; zero for l <= max(m, 2)):
//   X_l = (a z + b) X_{l-1} - c X_{l-2},   Y_l = (a z - b) Y_{l-1} - c Y_{l-2},
//   a = sqrt((2L+3)(2L+1)) (L+1) / q,  b = 2 m a / (L (L+1)),
//   c = sqrt((2L+3)/(2L-1)) (L+1) sqrt((L^2-m^2)(L^2-4)) / (L q),
//   q = sqrt(((L+1)^2-m^2)((L+1)^2-4))
// (the Wigner-d recurrence in l with m' = -+2, times N_l).  The integer
// products stay below 2^53 for l <= 8191.
__global__ void __launch_bounds__(256) rec_coefs_s2(int lmax, double* __restrict__ abc) {
  const int m = blockIdx.y;
  const int l = m + blockIdx.x * blockDim.x + threadIdx.x;
  if (l > lmax) return;
  const int ls = m > 2 ? m : 2;
  double a = 0.0, b = 0.0, c = 0.0;
  if (l > ls) {
    const double L = l - 1, L1 = l, M = m;
    const double q = sqrt((L1 - M) * (L1 + M) * (L1 - 2.0) * (L1 + 2.0));
    a = sqrt((2.0 * L + 3.0) * (2.0 * L + 1.0)) * L1 / q;
    b = a * 2.0 * M / (L * L1);
    c = sqrt((2.0 * L + 3.0) / (2.0 * L - 1.0)) * L1 *
        sqrt((L - M) * (L + M) * (L - 2.0) * (L + 2.0)) / (L * q);
  }
  double* o = abc + 3 * alm_idx(lmax, l, m);
  o[0] = a;
  o[1] = b;
  o[2] = c;
}

// X and Y at l = max(m, 2) on north ring k, as mantissas of a common scale
// 2^(600 s).  With c = cos(theta/2), s = sin(theta/2), K_m^2 = (2m)!/((m+2)!(m-2)!):
//   m >= 2: X = (-1)^m N_m K_m (cs)^(m-2) s^4,  Y = ... c^4
//           (s^4 = (1-z)^2/4, c^4 = (1+z)^2/4, cs = sin(theta)/2);
//   m = 1:  X = -N_2 sin(theta) (1-z)/2,  Y = N_2 sin(theta) (1+z)/2;
//   m = 0:  X = Y = N_2 (sqrt 6 / 4) sin^2(theta).
__device__ inline void pol_start(const PolArgs& A, int k, int m, double& x, double& y, int& s) {
  const double omz = A.omz[k];
  const double opz = 1.0 + A.R.z[k];
  if (m >= 2) {
    const double L = A.logc2[m] + (double)(m - 2) * A.R.logsin[k];
    const int sc = L >= 0.0 ? 0 : -(int)floor(-L / kLnBig);
    double base = exp(L - (double)sc * kLnBig);
    if (m & 1) base = -base;
    x = 0.25 * base * omz * omz;
    y = 0.25 * base * opz * opz;
    s = sc;
  } else {
    const double n2 = 0.63078313050504001;   // sqrt(5 / (4 pi))
    const double st = exp(A.R.logsin[k]);
    if (m == 0) {
      x = y = n2 * 0.61237243569579452 * omz * opz;   // sqrt(6)/4 sin^2
    } else {
      x = -0.5 * n2 * st * omz;
      y = 0.5 * n2 * st * opz;
    }
    s = 0;
  }
}

// One step of the X/Y recurrences of a ring pair (the start value at l = ls);
// returns 2 lambda+ and 2 lambda- weighted by the pair's scale (see
// rescale_pair: rescaling runs once per chunk of steps).
__device__ __forceinline__ void pol_adv(bool start, double a, double b, double c, double z,
                                        double& x1, double& x2, double& y1, double& y2,
                                        double wgt, double& tp, double& tm) {
  double x, y;
  if (start) {
    x = x1;
    y = y1;
  } else {
    const double az = a * z;
    x = fma(az + b, x1, -c * x2);
    y = fma(az - b, y1, -c * y2);
    x2 = x1;
    x1 = x;
    y2 = y1;
    y1 = y;
  }
  tp = (x + y) * wgt;
  tm = (x - y) * wgt;
}

__device__ __forceinline__ void pol_rescale(double& x1, double& x2, double& y1, double& y2,
                                            int& s, double& wgt) {
  if (s < 0 && fmax(fabs(x1), fabs(y1)) > kBig) {
    x1 *= kInvBig;
    x2 *= kInvBig;
    y1 *= kInvBig;
    y2 *= kInvBig;
    s += 1;
    wgt = scale_weight(s);
  }
}

// Spin-2 analysis partial sums: part[comp][chunk][idx(l, m)], comp 0 = E, 1 = B,
// of sum over the chunk's pairs of [2 lambda+ Q + 2i lambda- U] (E) and
// [2 lambda+ U - 2i lambda- Q] (B); F holds Q+, Q-, U+, U- at stride fs.
// The spin-2 analysis asks for PSX_AN2_WAVES waves per SIMD with
// kAn2Pairs ring pairs per thread (152 VGPRs at 4 pairs and 3 waves, no
// spills).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PSX_AN2_WAVES)))
legendre_analysis_s2(PolArgs A, const double2* __restrict__ F,
                                                            int64_t fs, double2* __restrict__ part) {
  __shared__ double red[kLChunk2][4][256];
  const int m = blockIdx.y;
  const int chunk = blockIdx.x;
  const int lmax = A.R.lmax;
  const int tid = threadIdx.x;
  const int ls = m > 2 ? m : 2;
  const double* Cm = A.abc + 3 * alm_idx(lmax, m, m);   // Cm[3 (l - m) + 0..2]
  const int64_t nalm = alm_idx(lmax, lmax + 1, lmax + 1);
  double z[kAn2Pairs], x1[kAn2Pairs], x2[kAn2Pairs];
  double y1[kAn2Pairs], y2[kAn2Pairs], wgt[kAn2Pairs];
  int sc[kAn2Pairs];
  double2 qp[kAn2Pairs], qm[kAn2Pairs], up[kAn2Pairs], um[kAn2Pairs];
#pragma unroll
  for (int r = 0; r < kAn2Pairs; ++r) {
    const int p = (chunk * kAn2Pairs + r) * 256 + tid;
    const bool live = p < A.R.npairs;
    const int pp = live ? p : 0;
    z[r] = A.R.z[pp];
    pol_start(A, pp, m, x1[r], y1[r], sc[r]);
    wgt[r] = scale_weight(sc[r]);
    x2[r] = y2[r] = 0.0;
    const int64_t o = (int64_t)m * A.R.npairs + pp;   // in bounds for every lane
    const double keep = live ? 1.0 : 0.0;
    qp[r] = F[o];
    qm[r] = F[fs + o];
    up[r] = F[2 * fs + o];
    um[r] = F[3 * fs + o];
    qp[r].x *= keep;
    qp[r].y *= keep;
    qm[r].x *= keep;
    qm[r].y *= keep;
    up[r].x *= keep;
    up[r].y *= keep;
    um[r].x *= keep;
    um[r].y *= keep;
  }
  const int o = tid >> 4, g = tid & 15;
  const int oj = o >> 2, oc = o & 3;
  double* const outp = (double*)(part + ((int64_t)(oc >> 1) * A.nchunk + chunk) * nalm);
  // the first round holds l = ls, the last may pass lmax: both take the
  // checked path, the rounds between carry no per-step branch
  auto round = [&](const int l0, const bool edge) {
#pragma unroll
    for (int j = 0; j < kLChunk2; ++j) {
      const int l = l0 + j;
      double er = 0.0, ei = 0.0, br = 0.0, bi = 0.0;
      if (!edge || (l <= lmax && l >= ls)) {
        double a = 0.0, b = 0.0, c = 0.0;
        if (!edge || l > ls) {
          const double* cf = Cm + 3 * (l - m);
          a = sload(cf);
          b = sload(cf + 1);
          c = sload(cf + 2);
        }
        const bool odd = j & 1;                    // l0 - m is a multiple of kLChunk2
#pragma unroll
        for (int r = 0; r < kAn2Pairs; ++r) {
          double tp, tm;
          pol_adv(edge && l == ls, a, b, c, z[r], x1[r], x2[r], y1[r], y2[r], wgt[r], tp, tm);
          // lambda+ pairs with the (-1)^{l+m} fold, lambda- with the other
          const double2 qa = odd ? qm[r] : qp[r], qb = odd ? qp[r] : qm[r];
          const double2 ua = odd ? um[r] : up[r], ub = odd ? up[r] : um[r];
          er = fma(tp, qa.x, fma(-tm, ub.y, er));
          ei = fma(tp, qa.y, fma(tm, ub.x, ei));
          br = fma(tp, ua.x, fma(tm, qb.y, br));
          bi = fma(tp, ua.y, fma(-tm, qb.x, bi));
        }
      }
      red[j][0][tid] = er;
      red[j][1][tid] = ei;
      red[j][2][tid] = br;
      red[j][3][tid] = bi;
    }
#pragma unroll
    for (int r = 0; r < kAn2Pairs; ++r)
      if (!edge || l0 + kLChunk2 > ls) pol_rescale(x1[r], x2[r], y1[r], y2[r], sc[r], wgt[r]);
    __syncthreads();
    {
      // 16 lanes per output, each summing a stride-16 column of the row
      double h[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double* row = &red[oj][oc][64 * k + g];
        h[k] = (row[0] + row[16]) + (row[32] + row[48]);
      }
      const double sum = row16_sum((h[0] + h[1]) + (h[2] + h[3]));
      const int l = l0 + oj;
      if (g == 0 && l <= lmax) outp[2 * alm_idx(lmax, l, m) + (oc & 1)] = sum;
    }
    __syncthreads();
  };
  int l0 = m;
  round(l0, true);
  for (l0 += kLChunk2; l0 + kLChunk2 - 1 <= lmax; l0 += kLChunk2) round(l0, false);
  if (l0 <= lmax) round(l0, true);
}

// Spin-2 synthesis: GQ[m][k], GU[m][k] = Q_m, U_m of every ring, kSyn2Pairs
// ring pairs per thread; lambda+ terms go to the part of the (-1)^{l+m}
// parity, lambda- terms to the other (north = even + odd, south = even - odd).
// a_E, a_B and the coefficients of kLChunk2 steps are read ahead with scalar
// loads.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PSX_SYN2_WAVES)))
legendre_synthesis_s2(PolArgs A,
                                                             const double2* __restrict__ aE,
                                                             const double2* __restrict__ aB,
                                                             double2* __restrict__ GQ,
                                                             double2* __restrict__ GU) {
  const int m = blockIdx.y;
  const int lmax = A.R.lmax;
  const int ls = m > 2 ? m : 2;
  const int64_t i0 = alm_idx(lmax, m, m);
  const double2* em = aE + i0;
  const double2* bm = aB + i0;
  const double* Cm = A.abc + 3 * i0;
  double z[kSyn2Pairs], x1[kSyn2Pairs], x2[kSyn2Pairs], y1[kSyn2Pairs], y2[kSyn2Pairs];
  double wgt[kSyn2Pairs];
  double2 qe[kSyn2Pairs], qo[kSyn2Pairs], ue[kSyn2Pairs], uo[kSyn2Pairs];
  int sc[kSyn2Pairs], pr[kSyn2Pairs];
#pragma unroll
  for (int r = 0; r < kSyn2Pairs; ++r) {
    const int p = (blockIdx.x * kSyn2Pairs + r) * blockDim.x + threadIdx.x;
    pr[r] = p;
    const int pp = p < A.R.npairs ? p : 0;
    z[r] = A.R.z[pp];
    pol_start(A, pp, m, x1[r], y1[r], sc[r]);
    wgt[r] = scale_weight(sc[r]);
    x2[r] = y2[r] = 0.0;
    qe[r] = qo[r] = ue[r] = uo[r] = make_double2(0.0, 0.0);
  }
  // the first round holds l = ls, the last may pass lmax: both take the
  // checked path, the rounds between load and step unconditionally
  auto round = [&](const int l0, const bool edge) {
    double2 e[kLChunk2], b[kLChunk2];
    double ca[kLChunk2], cb[kLChunk2], cc[kLChunk2];
#pragma unroll
    for (int j = 0; j < kLChunk2; ++j) {
      const int l = l0 + j;
      const bool in = !edge || (l >= ls && l <= lmax);
      e[j] = in ? sload(em + (l - m)) : make_double2(0.0, 0.0);
      b[j] = in ? sload(bm + (l - m)) : make_double2(0.0, 0.0);
      const double* cf = Cm + 3 * (l - m);
      const bool rec = !edge || (in && l > ls);
      ca[j] = rec ? sload(cf) : 0.0;
      cb[j] = rec ? sload(cf + 1) : 0.0;
      cc[j] = rec ? sload(cf + 2) : 0.0;
    }
#pragma unroll
    for (int j = 0; j < kLChunk2; ++j) {
      const int l = l0 + j;
      if (edge && l > lmax) break;
      if (edge && l < ls) continue;
      const bool odd = j & 1;                      // l0 - m is a multiple of kLChunk2
#pragma unroll
      for (int r = 0; r < kSyn2Pairs; ++r) {
        double tp, tm;
        pol_adv(edge && l == ls, ca[j], cb[j], cc[j], z[r], x1[r], x2[r], y1[r], y2[r], wgt[r],
                tp, tm);
        // lambda+ a_E and i lambda- a_B (Q); lambda+ a_B and -i lambda- a_E (U)
        const double2 pq = make_double2(tp * e[j].x, tp * e[j].y);
        const double2 mq = make_double2(-tm * b[j].y, tm * b[j].x);
        const double2 pu = make_double2(tp * b[j].x, tp * b[j].y);
        const double2 mu = make_double2(tm * e[j].y, -tm * e[j].x);
        const double2 q_e = odd ? mq : pq, q_o = odd ? pq : mq;
        const double2 u_e = odd ? mu : pu, u_o = odd ? pu : mu;
        qe[r].x += q_e.x;
        qe[r].y += q_e.y;
        qo[r].x += q_o.x;
        qo[r].y += q_o.y;
        ue[r].x += u_e.x;
        ue[r].y += u_e.y;
        uo[r].x += u_o.x;
        uo[r].y += u_o.y;
      }
    }
#pragma unroll
    for (int r = 0; r < kSyn2Pairs; ++r)
      if (!edge || l0 + kLChunk2 > ls) pol_rescale(x1[r], x2[r], y1[r], y2[r], sc[r], wgt[r]);
  };
  int l0 = m;
  round(l0, true);
  for (l0 += kLChunk2; l0 + kLChunk2 - 1 <= lmax; l0 += kLChunk2) round(l0, false);
  if (l0 <= lmax) round(l0, true);
#pragma unroll
  for (int r = 0; r < kSyn2Pairs; ++r) {
    const int p = pr[r];
    if (p >= A.R.npairs) continue;
    const int kn = p, ks = A.R.nr - 1 - p;
    const int64_t on = gidx(lmax, m, kn), os = gidx(lmax, m, ks);
    // the factor -1/2 turns 2 lambda+- sums into Q_m, U_m
    GQ[on] = make_double2(-0.5 * (qe[r].x + qo[r].x), -0.5 * (qe[r].y + qo[r].y));
    GU[on] = make_double2(-0.5 * (ue[r].x + uo[r].x), -0.5 * (ue[r].y + uo[r].y));
    if (ks != kn) {
      GQ[os] = make_double2(-0.5 * (qe[r].x - qo[r].x), -0.5 * (qe[r].y - qo[r].y));
      GU[os] = make_double2(-0.5 * (ue[r].x - uo[r].x), -0.5 * (ue[r].y - uo[r].y));
    }
  }
}

// ---- cross-spectra ----------------------------------------------------------
// C^{ij}_l = (2l+1)^-1 [Re(a^i_l0 a^j_l0*) + 2 sum_{m>=1} Re(a^i_lm a^j_lm*)]
// for every pair i <= j of NM maps (Healpix alm2cl as the reference applies it
// to mask alms, src/modecoupling.jl:197, src/workspace.jl:202).  Thread l of
// chunk c sums m in [c mchunk, (c+1) mchunk) (the a_lm of consecutive l are
// adjacent, so the loads coalesce); chunk partials part[c][pair][l] are then
// added in chunk order by cross_spectra_sum, so the result does not depend on
// scheduling.
template <int NM>
__global__ void __launch_bounds__(256) cross_spectra_partial(const double2* __restrict__ alms,
                                                             int64_t nalm, int lmax, int mchunk,
                                                             double* __restrict__ part) {
  constexpr int NP = NM * (NM + 1) / 2;
  const int l = (int)blockIdx.x * 256 + (int)threadIdx.x;
  const int c = blockIdx.y;
  if (l > lmax) return;
  double acc[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) acc[q] = 0.0;
  const int m0 = c * mchunk, m1 = min(m0 + mchunk, l + 1);
  for (int m = m0; m < m1; ++m) {
    const int64_t idx = alm_idx(lmax, l, m);
    const double w = m == 0 ? 1.0 : 2.0;
    double2 a[NM];
#pragma unroll
    for (int i = 0; i < NM; ++i) a[i] = alms[i * nalm + idx];
    int q = 0;
#pragma unroll
    for (int i = 0; i < NM; ++i)
#pragma unroll
      for (int j = i; j < NM; ++j, ++q)
        acc[q] = fma(w, fma(a[i].x, a[j].x, a[i].y * a[j].y), acc[q]);
  }
  const int64_t L = lmax + 1;
#pragma unroll
  for (int q = 0; q < NP; ++q) part[((int64_t)c * NP + q) * L + l] = acc[q];
}

// cl[i][j][l] = cl[j][i][l] = sum_c part[c][pair(i,j)][l] / (2l+1)
__global__ void __launch_bounds__(256) cross_spectra_sum(const double* __restrict__ part,
                                                         int nchunk, int nm, int lmax,
                                                         double* __restrict__ cl) {
  const int l = (int)blockIdx.x * 256 + (int)threadIdx.x;
  const int q = blockIdx.y;                      // pair index, i <= j in row order
  if (l > lmax) return;
  const int np = nm * (nm + 1) / 2;
  const int64_t L = lmax + 1;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c) s += part[((int64_t)c * np + q) * L + l];
  s /= 2.0 * l + 1.0;
  int i = 0, r = q;
  while (r >= nm - i) {
    r -= nm - i;
    ++i;
  }
  const int j = i + r;
  cl[((int64_t)i * nm + j) * L + l] = s;
  cl[((int64_t)j * nm + i) * L + l] = s;
}

__global__ void __launch_bounds__(256) residual(const double* __restrict__ a,
                                                const double* __restrict__ b, int64_t n,
                                                double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] - b[i];
}

// x[i] *= w[i]: a map times its mask, in place
__global__ void __launch_bounds__(256) apply_mask(double* __restrict__ x,
                                                  const double* __restrict__ w, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] *= w[i];
}

// The polar DFTs with LDS tables sized for the longest ring of the grid
// (4 nside - 4 pixels): 41 KB and 256 threads for nside <= 1024, 82 KB and
// 512 threads up to 2048.
void launch_polar_dft(int ns, hipStream_t s, RingArgs ra, const double* map, double2* X) {
  constexpr int kSmall = kMaxPolarRing / 2;
  if (4 * (ns - 1) <= kSmall)
    hipLaunchKernelGGL(polar_dft<kSmall>, dim3(2 * (ns - 1)), dim3(kSmall / 16), 0, s, ra, map, X,
                       ns);
  else
    hipLaunchKernelGGL(polar_dft<kMaxPolarRing>, dim3(2 * (ns - 1)), dim3(kMaxPolarRing / 16), 0,
                       s, ra, map, X, ns);
}

void launch_polar_idft(int ns, hipStream_t s, RingArgs ra, const double2* X, double* map) {
  constexpr int kSmall = kMaxPolarRing / 2;
  if (4 * (ns - 1) <= kSmall)
    hipLaunchKernelGGL(polar_idft<kSmall>, dim3(2 * (ns - 1)), dim3(kSmall / 16), 0, s, ra, X, map,
                       ns);
  else
    hipLaunchKernelGGL(polar_idft<kMaxPolarRing>, dim3(2 * (ns - 1)), dim3(kMaxPolarRing / 16), 0,
                       s, ra, X, map, ns);
}

// ---- plans ----------------------------------------------------------------
#define PSX_FFT(call)                                                            \
  do {                                                                           \
    rocfft_status s_ = (call);                                                   \
    if (s_ != rocfft_status_success)                                             \
      throw ::psx::Error{PSX_ERR_DEVICE, std::string(#call) + " failed (" +      \
                                             std::to_string((int)s_) + ")"};     \
  } while (0)

struct ShtPlan {
  Rings R;
  int lmax = 0;
  int64_t nalm = 0;
  int nchunk = 0;          // pair chunks of the spin-0 analysis
  int nchunk2 = 0;         // pair chunks of the spin-2 analysis
  // device
  int* dn = nullptr;
  int64_t* dfirst = nullptr;
  int64_t* dxoff = nullptr;
  double *dz = nullptr, *dlogsin = nullptr, *dphi0 = nullptr, *dlogc = nullptr;
  double2* dAB = nullptr;
  // spin 2, built on first use
  double *domz = nullptr, *dabc = nullptr, *dlogc2 = nullptr;
  rocfft_plan fwd = nullptr, inv = nullptr;
  rocfft_execution_info info = nullptr;
  void* work = nullptr;
  size_t work_bytes = 0;
  RingArgs ra() const {
    return RingArgs{dn, dfirst, dxoff, dz, dlogsin, dphi0, R.nr, R.npairs, lmax};
  }
};

std::mutex g_fft_mu;
bool g_fft_ready = false;
std::map<std::tuple<void*, int, int>, ShtPlan*> g_plans;

// Spin-2 tables of a plan: 1 - z per ring (long double, exact near the
// pole), the recurrence coefficients and log(N_m K_m) - (m - 2) ln 2 with
// log K_m = (1/2) sum_{j=2}^{m-1} log((2j+2)(2j+1) / ((j+3)(j-1))).
void build_pol_tables(Ctx* c, ShtPlan* P) {
  const Rings& R = P->R;
  const int lmax = P->lmax;
  const long double nsl = R.nside;
  std::vector<double> omz(R.nr), logc2(lmax + 1, 0.0);
  for (int k = 0; k < R.nr; ++k) {
    const int i = k + 1;
    long double v;
    if (i < R.nside) v = (long double)i * i / (3.0L * nsl * nsl);
    else if (i <= 3 * R.nside) v = 1.0L - (4.0L / 3.0L - 2.0L * i / (3.0L * nsl));
    else v = 1.0L + (1.0L - (long double)(4 * R.nside - i) * (4 * R.nside - i) / (3.0L * nsl * nsl));
    omz[k] = (double)v;
  }
  long double lk = 0.0L;
  for (int m = 2; m <= lmax; ++m) {
    if (m > 2) {
      const long double j = m - 1;
      lk += 0.5L * logl((2.0L * j + 2.0L) * (2.0L * j + 1.0L) / ((j + 3.0L) * (j - 1.0L)));
    }
    logc2[m] = (double)(0.5L * logl((2.0L * m + 1.0L) / (4.0L * (long double)kPi)) + lk -
                        (long double)(m - 2) * logl(2.0L));
  }
  P->domz = (double*)ctx_alloc(c, sizeof(double) * R.nr);
  P->dlogc2 = (double*)ctx_alloc(c, sizeof(double) * (lmax + 1));
  P->dabc = (double*)ctx_alloc(c, 3 * sizeof(double) * P->nalm);
  PSX_HIP(hipMemcpy(P->domz, omz.data(), sizeof(double) * R.nr, hipMemcpyHostToDevice));
  PSX_HIP(hipMemcpy(P->dlogc2, logc2.data(), sizeof(double) * (lmax + 1), hipMemcpyHostToDevice));
  dim3 grid((unsigned)((lmax + 1 + 255) / 256), (unsigned)(lmax + 1));
  hipLaunchKernelGGL(rec_coefs_s2, grid, dim3(256), 0, c->stream, lmax, P->dabc);
  PSX_HIP(hipGetLastError());
  PSX_HIP(hipStreamSynchronize(c->stream));
}

ShtPlan* get_plan(Ctx* c, int nside, int lmax, bool pol = false) {
  std::lock_guard<std::mutex> lk(g_fft_mu);
  auto key = std::make_tuple((void*)c, nside, lmax);
  auto it = g_plans.find(key);
  if (it != g_plans.end()) {
    if (pol && !it->second->dabc) build_pol_tables(c, it->second);
    return it->second;
  }
  if (!g_fft_ready) {
    PSX_FFT(rocfft_setup());
    g_fft_ready = true;
  }
  auto* P = new ShtPlan();
  P->R = make_rings(nside);
  P->lmax = lmax;
  P->nalm = alm_idx(lmax, lmax + 1, lmax + 1);
  P->nchunk = (P->R.npairs + 256 * kPairsPerThread - 1) / (256 * kPairsPerThread);
  P->nchunk2 = (P->R.npairs + 256 * kAn2Pairs - 1) / (256 * kAn2Pairs);
  const Rings& R = P->R;
  P->dn = (int*)ctx_alloc(c, sizeof(int) * R.nr);
  P->dfirst = (int64_t*)ctx_alloc(c, sizeof(int64_t) * R.nr);
  P->dxoff = (int64_t*)ctx_alloc(c, sizeof(int64_t) * R.nr);
  P->dz = (double*)ctx_alloc(c, sizeof(double) * R.nr);
  P->dlogsin = (double*)ctx_alloc(c, sizeof(double) * R.nr);
  P->dphi0 = (double*)ctx_alloc(c, sizeof(double) * R.nr);
  P->dlogc = (double*)ctx_alloc(c, sizeof(double) * (lmax + 1));
  P->dAB = (double2*)ctx_alloc(c, sizeof(double2) * P->nalm);
  PSX_HIP(hipMemcpy(P->dn, R.n.data(), sizeof(int) * R.nr, hipMemcpyHostToDevice));
  PSX_HIP(hipMemcpy(P->dfirst, R.first.data(), sizeof(int64_t) * R.nr, hipMemcpyHostToDevice));
  PSX_HIP(hipMemcpy(P->dxoff, R.xoff.data(), sizeof(int64_t) * R.nr, hipMemcpyHostToDevice));
  PSX_HIP(hipMemcpy(P->dz, R.z.data(), sizeof(double) * R.nr, hipMemcpyHostToDevice));
  PSX_HIP(hipMemcpy(P->dlogsin, R.logsin.data(), sizeof(double) * R.nr, hipMemcpyHostToDevice));
  PSX_HIP(hipMemcpy(P->dphi0, R.phi0.data(), sizeof(double) * R.nr, hipMemcpyHostToDevice));
  std::vector<double> logc(lmax + 1);
  long double acc = 0.0L;
  for (int m = 0; m <= lmax; ++m) {
    if (m > 0) acc += logl((2.0L * m - 1.0L) / (2.0L * m));
    logc[m] = (double)(0.5L * (logl((2.0L * m + 1.0L) / (4.0L * (long double)kPi)) + acc));
  }
  PSX_HIP(hipMemcpy(P->dlogc, logc.data(), sizeof(double) * (lmax + 1), hipMemcpyHostToDevice));
  {
    dim3 grid((unsigned)((lmax + 1 + 255) / 256), (unsigned)(lmax + 1));
    hipLaunchKernelGGL(rec_coefs, grid, dim3(256), 0, c->stream, lmax, P->dAB);
    PSX_HIP(hipGetLastError());
  }
  // equatorial rings: 2 nside + 1 transforms of length 4 nside
  const size_t len = (size_t)4 * nside, batch = (size_t)2 * nside + 1;
  const size_t hl = len / 2 + 1;
  {
    rocfft_plan_description d = nullptr;
    PSX_FFT(rocfft_plan_description_create(&d));
    const size_t one = 1;
    PSX_FFT(rocfft_plan_description_set_data_layout(d, rocfft_array_type_real,
                                                    rocfft_array_type_hermitian_interleaved,
                                                    nullptr, nullptr, 1, &one, len, 1, &one, hl));
    PSX_FFT(rocfft_plan_create(&P->fwd, rocfft_placement_notinplace,
                               rocfft_transform_type_real_forward, rocfft_precision_double, 1,
                               &len, batch, d));
    PSX_FFT(rocfft_plan_description_destroy(d));
  }
  {
    rocfft_plan_description d = nullptr;
    PSX_FFT(rocfft_plan_description_create(&d));
    const size_t one = 1;
    PSX_FFT(rocfft_plan_description_set_data_layout(d, rocfft_array_type_hermitian_interleaved,
                                                    rocfft_array_type_real, nullptr, nullptr, 1,
                                                    &one, hl, 1, &one, len));
    PSX_FFT(rocfft_plan_create(&P->inv, rocfft_placement_notinplace,
                               rocfft_transform_type_real_inverse, rocfft_precision_double, 1,
                               &len, batch, d));
    PSX_FFT(rocfft_plan_description_destroy(d));
  }
  size_t w1 = 0, w2 = 0;
  PSX_FFT(rocfft_plan_get_work_buffer_size(P->fwd, &w1));
  PSX_FFT(rocfft_plan_get_work_buffer_size(P->inv, &w2));
  P->work_bytes = std::max(w1, w2);
  if (P->work_bytes) P->work = ctx_alloc(c, P->work_bytes);
  PSX_FFT(rocfft_execution_info_create(&P->info));
  if (P->work) PSX_FFT(rocfft_execution_info_set_work_buffer(P->info, P->work, P->work_bytes));
  PSX_FFT(rocfft_execution_info_set_stream(P->info, (void*)c->stream));
  PSX_HIP(hipStreamSynchronize(c->stream));
  g_plans[key] = P;
  if (pol) build_pol_tables(c, P);
  return P;
}

}  // namespace

// Destroy the transform plans of a context that is being destroyed (their
// device buffers go with the context's allocations).
void release_sht_plans(Ctx* c) {
  std::lock_guard<std::mutex> lk(g_fft_mu);
  for (auto it = g_plans.begin(); it != g_plans.end();) {
    if (std::get<0>(it->first) == (void*)c) {
      ShtPlan* P = it->second;
      if (P->info) (void)rocfft_execution_info_destroy(P->info);
      if (P->fwd) (void)rocfft_plan_destroy(P->fwd);
      if (P->inv) (void)rocfft_plan_destroy(P->inv);
      delete P;
      it = g_plans.erase(it);
    } else {
      ++it;
    }
  }
  // the library's setup is paired with its cleanup once no plan is left, so
  // nothing of it is left for the process's exit to tear down
  if (g_plans.empty() && g_fft_ready) {
    (void)rocfft_cleanup();
    g_fft_ready = false;
  }
}

namespace {

// Device working set for one map.
struct ShtBufs {
  double2 *X = nullptr, *Fp = nullptr, *Fm = nullptr, *part = nullptr, *G = nullptr;
  double *syn = nullptr, *res = nullptr;
};

size_t sht_work_bytes(const ShtPlan& P) {
  const Rings& R = P.R;
  const size_t fl = (size_t)(P.lmax + 1) * R.npairs;
  return sizeof(double2) * ((size_t)R.xlen + 2 * fl + (size_t)P.nchunk * P.nalm +
                            (size_t)g_elems(P.lmax, R.nr)) +
         2 * sizeof(double) * (size_t)R.npix + 8 * 256;  // 7 buffers, 256-aligned
}

ShtBufs carve(const ShtPlan& P, void* mem) {
  const Rings& R = P.R;
  char* p = (char*)mem;
  auto take = [&](size_t b) {
    char* q = p;
    p += (b + 255) & ~size_t(255);
    return q;
  };
  const size_t fl = (size_t)(P.lmax + 1) * R.npairs;
  ShtBufs B;
  B.X = (double2*)take(sizeof(double2) * R.xlen);
  B.Fp = (double2*)take(sizeof(double2) * fl);
  B.Fm = (double2*)take(sizeof(double2) * fl);
  B.part = (double2*)take(sizeof(double2) * P.nchunk * P.nalm);
  B.G = (double2*)take(sizeof(double2) * g_elems(P.lmax, R.nr));
  B.syn = (double*)take(sizeof(double) * R.npix);
  B.res = (double*)take(sizeof(double) * R.npix);
  return B;
}

LegArgs leg(const ShtPlan& P) { return LegArgs{P.ra(), P.dAB, P.dlogc, P.nchunk}; }

// Ring Fourier coefficients of one map, folded into B.Fp, B.Fm.
void fold_in(Ctx* c, ShtPlan& P, const ShtBufs& B, const double* dmap) {
  const Rings& R = P.R;
  const int ns = R.nside;
  hipStream_t s = c->stream;
  if (ns > 1) {
    launch_polar_dft(ns, s, P.ra(), dmap, B.X);
    PSX_HIP(hipGetLastError());
  }
  {
    void* in[1] = {(void*)(dmap + R.first[ns - 1])};
    void* out[1] = {(void*)(B.X + R.xoff[ns - 1])};
    PSX_FFT(rocfft_execute(P.fwd, in, out, P.info));
  }
  hipLaunchKernelGGL(fold_analysis, dim3((unsigned)((R.npairs + kFaPairs - 1) / kFaPairs),
                                         (unsigned)((P.lmax + kFaM) / kFaM)),
                     dim3(256), 0, s, P.ra(), B.X, B.Fp, B.Fm);
  PSX_HIP(hipGetLastError());
}

template <int NM>
void launch_analysis(ShtPlan& P, hipStream_t s, const ShtBufs* B) {
  MapPtrs<NM> M{};
  for (int q = 0; q < NM; ++q) {
    M.in[q] = B[q].Fp;
    M.in2[q] = B[q].Fm;
    M.out[q] = B[q].part;
  }
  hipLaunchKernelGGL(legendre_analysis<NM>, dim3((unsigned)P.nchunk, (unsigned)(P.lmax + 1)),
                     dim3(256), 0, s, leg(P), M);
  PSX_HIP(hipGetLastError());
}

template <int NM>
void launch_synthesis(ShtPlan& P, hipStream_t s, const ShtBufs* B, const double2* const* alms) {
  MapPtrs<NM> M{};
  for (int q = 0; q < NM; ++q) {
    M.in[q] = alms[q];
    M.out[q] = B[q].G;
  }
  constexpr int SP = syn_pairs(NM);
  hipLaunchKernelGGL(legendre_synthesis<NM>,
                     dim3((unsigned)((P.R.npairs + 256 * SP - 1) / (256 * SP)),
                          (unsigned)(P.lmax + 1)),
                     dim3(256), 0, s, leg(P), M);
  PSX_HIP(hipGetLastError());
}

// maps per analysis pass (the analysis holds each map's F+- of its ring pairs
// in registers and reduces 2 values per map and l through LDS)
#ifndef PSX_AN_GROUP
#define PSX_AN_GROUP 2
#endif

// alms[q] (+)= A(maps[q]) for a group of nm <= kMaxGroup maps: Legendre passes
// of up to PSX_AN_GROUP maps (the recurrence shared), ring transforms per map.
// The maps of each analysis launch are folded just before it; `before`, if
// given, is called with q before map q is first read.
void analysis_group(Ctx* c, ShtPlan& P, const ShtBufs* B, int nm, const double* const* maps,
                    double2* const* alms, bool accumulate,
                    const std::function<void(int)>* before = nullptr) {
  PSX_REQUIRE(nm >= 1 && nm <= kMaxGroup, PSX_ERR_ARGUMENT, "1..4 maps per transform group");
  hipStream_t s = c->stream;
  for (int q0 = 0; q0 < nm; q0 += PSX_AN_GROUP) {
    const int k = std::min(PSX_AN_GROUP, nm - q0);
    for (int q = q0; q < q0 + k; ++q) {
      if (before) (*before)(q);
      fold_in(c, P, B[q], maps[q]);
    }
    switch (k) {
      case 1: launch_analysis<1>(P, s, B + q0); break;
      case 2: launch_analysis<2>(P, s, B + q0); break;
      case 3: launch_analysis<3>(P, s, B + q0); break;
      default: launch_analysis<4>(P, s, B + q0); break;
    }
  }
  const double w = 4.0 * kPi / (double)P.R.npix;
  for (int q = 0; q < nm; ++q) {
    hipLaunchKernelGGL(sum_partials, dim3((unsigned)((P.nalm + 255) / 256)), dim3(256), 0, s,
                       B[q].part, P.nchunk, P.nalm, w, accumulate ? 1 : 0, alms[q]);
    PSX_HIP(hipGetLastError());
  }
}

// maps[q] = S(alms[q]) for a group of nm <= kMaxGroup coefficient sets
void synthesis_group(Ctx* c, ShtPlan& P, const ShtBufs* B, int nm, const double2* const* alms,
                     double* const* maps) {
  PSX_REQUIRE(nm >= 1 && nm <= kMaxGroup, PSX_ERR_ARGUMENT, "1..4 maps per transform group");
  const Rings& R = P.R;
  const int ns = R.nside;
  hipStream_t s = c->stream;
  switch (nm) {
    case 1: launch_synthesis<1>(P, s, B, alms); break;
    case 2: launch_synthesis<2>(P, s, B, alms); break;
    case 3: launch_synthesis<3>(P, s, B, alms); break;
    default: launch_synthesis<4>(P, s, B, alms); break;
  }
  for (int q = 0; q < nm; ++q) {
    launch_fold_synthesis(R, s, P.ra(), B[q].G, B[q].X);
    {
      void* in[1] = {(void*)(B[q].X + R.xoff[ns - 1])};
      void* out[1] = {(void*)(maps[q] + R.first[ns - 1])};
      PSX_FFT(rocfft_execute(P.inv, in, out, P.info));
    }
    if (ns > 1) {
      launch_polar_idft(ns, s, P.ra(), B[q].X, maps[q]);
      PSX_HIP(hipGetLastError());
    }
  }
}

// Device working set for one polarised (Q, U) pair.
struct PolBufs {
  double2 *X[2] = {nullptr, nullptr}, *F = nullptr, *part = nullptr;
  double2* G[2] = {nullptr, nullptr};
  double *syn = nullptr, *res = nullptr;   // Q then U, npix each
  int64_t fs = 0;
};

size_t pol_work_bytes(const ShtPlan& P) {
  const Rings& R = P.R;
  const size_t fl = (size_t)(P.lmax + 1) * R.npairs;
  return sizeof(double2) * (2 * (size_t)R.xlen + 4 * fl + 2 * (size_t)P.nchunk2 * P.nalm +
                            2 * (size_t)g_elems(P.lmax, R.nr)) +
         4 * sizeof(double) * (size_t)R.npix + 8 * 256;  // 7 buffers, 256-aligned
}

PolBufs carve_pol(const ShtPlan& P, void* mem) {
  const Rings& R = P.R;
  char* p = (char*)mem;
  auto take = [&](size_t b) {
    char* q = p;
    p += (b + 255) & ~size_t(255);
    return q;
  };
  const size_t fl = (size_t)(P.lmax + 1) * R.npairs;
  PolBufs B;
  B.X[0] = (double2*)take(sizeof(double2) * R.xlen);
  B.X[1] = (double2*)take(sizeof(double2) * R.xlen);
  B.F = (double2*)take(4 * sizeof(double2) * fl);
  B.fs = (int64_t)fl;
  B.part = (double2*)take(2 * sizeof(double2) * P.nchunk2 * P.nalm);
  B.G[0] = (double2*)take(sizeof(double2) * g_elems(P.lmax, R.nr));
  B.G[1] = (double2*)take(sizeof(double2) * g_elems(P.lmax, R.nr));
  B.syn = (double*)take(2 * sizeof(double) * R.npix);
  B.res = (double*)take(2 * sizeof(double) * R.npix);
  return B;
}

PolArgs pleg(const ShtPlan& P) { return PolArgs{P.ra(), P.domz, P.dabc, P.dlogc2, P.nchunk2}; }

// Ring Fourier coefficients of one real map, folded into F+ and F-.
void ring_fold(Ctx* c, ShtPlan& P, const double* dmap, double2* X, double2* Fp, double2* Fm) {
  const Rings& R = P.R;
  const int ns = R.nside;
  hipStream_t s = c->stream;
  if (ns > 1) {
    launch_polar_dft(ns, s, P.ra(), dmap, X);
    PSX_HIP(hipGetLastError());
  }
  {
    void* in[1] = {(void*)(dmap + R.first[ns - 1])};
    void* out[1] = {(void*)(X + R.xoff[ns - 1])};
    PSX_FFT(rocfft_execute(P.fwd, in, out, P.info));
  }
  hipLaunchKernelGGL(fold_analysis, dim3((unsigned)((R.npairs + kFaPairs - 1) / kFaPairs),
                                         (unsigned)((P.lmax + kFaM) / kFaM)),
                     dim3(256), 0, s, P.ra(), X, Fp, Fm);
  PSX_HIP(hipGetLastError());
}

// Ring spectra G -> map (inverse fold, c2r rocFFT, polar inverse DFT).
void ring_unfold(Ctx* c, ShtPlan& P, const double2* G, double2* X, double* dmap) {
  const Rings& R = P.R;
  const int ns = R.nside;
  hipStream_t s = c->stream;
  launch_fold_synthesis(R, s, P.ra(), G, X);
  {
    void* in[1] = {(void*)(X + R.xoff[ns - 1])};
    void* out[1] = {(void*)(dmap + R.first[ns - 1])};
    PSX_FFT(rocfft_execute(P.inv, in, out, P.info));
  }
  if (ns > 1) {
    launch_polar_idft(ns, s, P.ra(), X, dmap);
    PSX_HIP(hipGetLastError());
  }
}

// (a_E, a_B) (+)= A2(Q, U)
void analysis_pol(Ctx* c, ShtPlan& P, const PolBufs& B, const double* dQ, const double* dU,
                  double2* aE, double2* aB, bool accumulate) {
  const Rings& R = P.R;
  hipStream_t s = c->stream;
  ring_fold(c, P, dQ, B.X[0], B.F, B.F + B.fs);
  ring_fold(c, P, dU, B.X[1], B.F + 2 * B.fs, B.F + 3 * B.fs);
  hipLaunchKernelGGL(legendre_analysis_s2, dim3((unsigned)P.nchunk2, (unsigned)(P.lmax + 1)),
                     dim3(256), 0, s, pleg(P), (const double2*)B.F, B.fs, B.part);
  PSX_HIP(hipGetLastError());
  // the kernel sums 2 lambda+- terms; a_E, a_B carry -w
  const double w = -0.5 * 4.0 * kPi / (double)R.npix;
  const unsigned nb = (unsigned)((P.nalm + 255) / 256);
  hipLaunchKernelGGL(sum_partials, dim3(nb), dim3(256), 0, s, (const double2*)B.part, P.nchunk2,
                     P.nalm, w, accumulate ? 1 : 0, aE);
  PSX_HIP(hipGetLastError());
  hipLaunchKernelGGL(sum_partials, dim3(nb), dim3(256), 0, s,
                     (const double2*)(B.part + (int64_t)P.nchunk2 * P.nalm), P.nchunk2, P.nalm, w,
                     accumulate ? 1 : 0, aB);
  PSX_HIP(hipGetLastError());
}

// (Q, U) = S2(a_E, a_B)
void synthesis_pol(Ctx* c, ShtPlan& P, const PolBufs& B, const double2* aE, const double2* aB,
                   double* dQ, double* dU) {
  const Rings& R = P.R;
  hipLaunchKernelGGL(legendre_synthesis_s2,
                     dim3((unsigned)((R.npairs + 256 * kSyn2Pairs - 1) / (256 * kSyn2Pairs)),
                          (unsigned)(P.lmax + 1)),
                     dim3(256), 0, c->stream, pleg(P), aE, aB, B.G[0], B.G[1]);
  PSX_HIP(hipGetLastError());
  ring_unfold(c, P, B.G[0], B.X[0], dQ);
  ring_unfold(c, P, B.G[1], B.X[1], dU);
}

void check_sht_args(int nside, int lmax) {
  // polar rings have at most 4 nside - 4 pixels: the DFT kernels' LDS
  // tables hold rings of up to kMaxPolarRing pixels
  PSX_REQUIRE(nside >= 1 && 4 * (nside - 1) <= kMaxPolarRing, PSX_ERR_ARGUMENT,
              "nside must be 1..2048");
  PSX_REQUIRE((nside & (nside - 1)) == 0, PSX_ERR_ARGUMENT, "nside must be a power of two");
  PSX_REQUIRE(lmax >= 0 && lmax <= 4 * nside && lmax <= 8191, PSX_ERR_ARGUMENT,
              "lmax must be 0..min(4 nside, 8191)");
}

}  // namespace

// map2alm of n device maps (any addresses) into device alms (complex
// interleaved), with niter refinements, on the context stream.  Maps go in
// groups of up to kMaxGroup that share their Legendre passes and iterate in
// lockstep; every result equals the one-map transform bit for bit.
// sync = false: return once everything is enqueued on the context stream
void map2alm_ptrs(Ctx* c, int nside, int lmax, int niter, int n, const double* const* maps,
                  double* const* alms, const std::function<void(int)>* before = nullptr,
                  bool sync = true) {
  check_sht_args(nside, lmax);
  ShtPlan& P = *get_plan(c, nside, lmax);
  const size_t per = (sht_work_bytes(P) + 255) & ~size_t(255);
  const int gmax = std::min(n, kMaxGroup);
  WorkBuf wk(c, per * (size_t)std::max(gmax, 1));
  ShtBufs B[kMaxGroup];
  for (int q = 0; q < gmax; ++q) B[q] = carve(P, (char*)wk.base() + q * per);
  const int64_t npix = P.R.npix;
  for (int g0 = 0; g0 < n; g0 += kMaxGroup) {
    const int nm = std::min(kMaxGroup, n - g0);
    const double* in[kMaxGroup];
    double2* alm[kMaxGroup];
    const double2* calm[kMaxGroup];
    const double* res[kMaxGroup];
    double* syn[kMaxGroup];
    for (int q = 0; q < nm; ++q) {
      in[q] = maps[g0 + q];
      alm[q] = (double2*)alms[g0 + q];
      calm[q] = alm[q];
      res[q] = B[q].res;
      syn[q] = B[q].syn;
    }
    if (before) {
      const std::function<void(int)> bq = [&](int q) { (*before)(g0 + q); };
      analysis_group(c, P, B, nm, in, alm, false, &bq);
    } else {
      analysis_group(c, P, B, nm, in, alm, false);
    }
    for (int it = 0; it < niter; ++it) {
      synthesis_group(c, P, B, nm, calm, syn);
      for (int q = 0; q < nm; ++q) {
        hipLaunchKernelGGL(residual, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0,
                           c->stream, in[q], (const double*)B[q].syn, npix, B[q].res);
        PSX_HIP(hipGetLastError());
      }
      analysis_group(c, P, B, nm, res, alm, true);
    }
  }
  if (sync) PSX_HIP(hipStreamSynchronize(c->stream));
}

// map2alm of nmaps consecutive device maps into consecutive alm sets
void map2alm_dev(Ctx* c, int nside, int lmax, int niter, int nmaps, const double* dmaps,
                 double* dalms, const std::function<void(int)>* before) {
  check_sht_args(nside, lmax);
  const int64_t npix = 12LL * nside * nside;
  const int64_t nalm = alm_idx(lmax, lmax + 1, lmax + 1);
  std::vector<const double*> m(nmaps);
  std::vector<double*> a(nmaps);
  for (int i = 0; i < nmaps; ++i) {
    m[i] = dmaps + i * npix;
    a[i] = dalms + 2 * i * nalm;
  }
  map2alm_ptrs(c, nside, lmax, niter, nmaps, m.data(), a.data(), before);
}

void alm2map_dev(Ctx* c, int nside, int lmax, int nmaps, const double* dalms, double* dmaps) {
  check_sht_args(nside, lmax);
  ShtPlan& P = *get_plan(c, nside, lmax);
  const size_t per = (sht_work_bytes(P) + 255) & ~size_t(255);
  const int gmax = std::min(nmaps, kMaxGroup);
  WorkBuf wk(c, per * (size_t)std::max(gmax, 1));
  ShtBufs B[kMaxGroup];
  for (int q = 0; q < gmax; ++q) B[q] = carve(P, (char*)wk.base() + q * per);
  for (int g0 = 0; g0 < nmaps; g0 += kMaxGroup) {
    const int nm = std::min(kMaxGroup, nmaps - g0);
    const double2* alm[kMaxGroup];
    double* map[kMaxGroup];
    for (int q = 0; q < nm; ++q) {
      alm[q] = (const double2*)dalms + (g0 + q) * P.nalm;
      map[q] = dmaps + (g0 + q) * P.R.npix;
    }
    synthesis_group(c, P, B, nm, alm, map);
  }
  PSX_HIP(hipStreamSynchronize(c->stream));
}

// Polarised map2alm: per map (I, Q, U) contiguous, npix each, into (T, E, B)
// alms, nalm complex each; niter refinements of both the spin-0 and the
// spin-2 parts.
void map2alm_pol_dev(Ctx* c, int nside, int lmax, int niter, int nmaps, const double* diqu,
                     double* dteb, bool spin0, const std::function<void()>* between) {
  check_sht_args(nside, lmax);
  ShtPlan& P = *get_plan(c, nside, lmax, true);
  const int64_t npix = P.R.npix, nalm = P.nalm;
  if (spin0) {
    std::vector<const double*> m(nmaps);
    std::vector<double*> a(nmaps);
    for (int i = 0; i < nmaps; ++i) {
      m[i] = diqu + 3 * i * npix;
      a[i] = dteb + 6 * i * nalm;
    }
    // with a `between` step the spin-0 work is left running meanwhile (the
    // work area it held is reused below only in stream order)
    map2alm_ptrs(c, nside, lmax, niter, nmaps, m.data(), a.data(), nullptr, between == nullptr);
  }
  if (between) (*between)();
  WorkBuf wk(c, pol_work_bytes(P));
  const PolBufs B = carve_pol(P, wk.base());
  for (int i = 0; i < nmaps; ++i) {
    const double* Q = diqu + (3 * i + 1) * npix;
    const double* U = Q + npix;
    double2* aE = (double2*)dteb + (3 * i + 1) * nalm;
    double2* aB = aE + nalm;
    analysis_pol(c, P, B, Q, U, aE, aB, false);
    for (int it = 0; it < niter; ++it) {
      synthesis_pol(c, P, B, aE, aB, B.syn, B.syn + npix);
      // Q and U are adjacent in the input, as are the two syntheses
      hipLaunchKernelGGL(residual, dim3((unsigned)((2 * npix + 255) / 256)), dim3(256), 0,
                         c->stream, Q, (const double*)B.syn, 2 * npix, B.res);
      PSX_HIP(hipGetLastError());
      analysis_pol(c, P, B, B.res, B.res + npix, aE, aB, true);
    }
  }
  PSX_HIP(hipStreamSynchronize(c->stream));
}

void alm2map_pol_dev(Ctx* c, int nside, int lmax, int nmaps, const double* dteb, double* diqu) {
  check_sht_args(nside, lmax);
  ShtPlan& P = *get_plan(c, nside, lmax, true);
  const int64_t npix = P.R.npix, nalm = P.nalm;
  for (int i = 0; i < nmaps; ++i)
    alm2map_dev(c, nside, lmax, 1, dteb + 6 * i * nalm, diqu + 3 * i * npix);
  WorkBuf wk(c, pol_work_bytes(P));
  const PolBufs B = carve_pol(P, wk.base());
  for (int i = 0; i < nmaps; ++i) {
    const double2* aE = (const double2*)dteb + (3 * i + 1) * nalm;
    double* Q = diqu + (3 * i + 1) * npix;
    synthesis_pol(c, P, B, aE, aE + nalm, Q, Q + npix);
  }
  PSX_HIP(hipStreamSynchronize(c->stream));
}

constexpr int kMaxSpectraMaps = 8;

// All cross-spectra of nmaps device alm sets (nalm complex each) into
// cl[nmaps][nmaps][lmax+1], on the context stream.
void cross_spectra_dev(Ctx* c, int lmax, int nmaps, const double* dalms, double* dcl) {
  PSX_REQUIRE(nmaps >= 1 && nmaps <= kMaxSpectraMaps, PSX_ERR_ARGUMENT,
              "cross-spectra of 1..8 maps per call");
  const int64_t nalm = alm_idx(lmax, lmax + 1, lmax + 1);
  const int np = nmaps * (nmaps + 1) / 2;
  const int nchunk = std::min(64, lmax + 1);
  const int mchunk = (lmax + 1 + nchunk - 1) / nchunk;
  WorkBuf wk(c, sizeof(double) * (size_t)nchunk * np * (lmax + 1));
  double* part = (double*)wk.base();
  {
    const dim3 grid((unsigned)((lmax + 1 + 255) / 256), (unsigned)nchunk);
    const double2* a = (const double2*)dalms;
    hipStream_t s = c->stream;
    switch (nmaps) {
#define PSX_CS(NMV)                                                                       \
  case NMV:                                                                               \
    hipLaunchKernelGGL(cross_spectra_partial<NMV>, grid, dim3(256), 0, s, a, nalm, lmax, \
                       mchunk, part);                                                     \
    break;
      PSX_CS(1) PSX_CS(2) PSX_CS(3) PSX_CS(4) PSX_CS(5) PSX_CS(6) PSX_CS(7) PSX_CS(8)
#undef PSX_CS
    }
    PSX_HIP(hipGetLastError());
    hipLaunchKernelGGL(cross_spectra_sum, dim3((unsigned)((lmax + 1 + 255) / 256), (unsigned)np),
                       dim3(256), 0, s, (const double*)part, nchunk, nmaps, lmax, dcl);
    PSX_HIP(hipGetLastError());
    PSX_HIP(hipStreamSynchronize(s));
  }
}

}  // namespace psx

using namespace psx;

extern "C" {

int psx_map2alm(psx_ctx* h, int nside, int lmax, int niter, int nmaps, const double* maps,
                double* alms) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  return guarded(c, [&] {
    PSX_REQUIRE(maps && alms && nmaps >= 1 && niter >= 0, PSX_ERR_ARGUMENT, "bad arguments");
    check_sht_args(nside, lmax);
    PSX_HIP(hipSetDevice(c->device));
    const int64_t npix = 12LL * nside * nside;
    const int64_t nalm = alm_idx(lmax, lmax + 1, lmax + 1);
    const size_t mb = sizeof(double) * npix * nmaps, ab = 2 * sizeof(double) * nalm * nmaps;
    XferBuf x(c, xfer_bytes({mb, ab}));
    double* dm = x.take(mb);
    double* da = x.take(ab);
    h2d_matrix(c, dm, npix, maps, npix, npix, nmaps);
    map2alm_dev(c, nside, lmax, niter, nmaps, dm, da);
    d2h_matrix(c, alms, 2 * nalm, da, 2 * nalm, 2 * nalm, nmaps);
  });
}

int psx_alm2map(psx_ctx* h, int nside, int lmax, int nmaps, const double* alms, double* maps) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  return guarded(c, [&] {
    PSX_REQUIRE(maps && alms && nmaps >= 1, PSX_ERR_ARGUMENT, "bad arguments");
    check_sht_args(nside, lmax);
    PSX_HIP(hipSetDevice(c->device));
    const int64_t npix = 12LL * nside * nside;
    const int64_t nalm = alm_idx(lmax, lmax + 1, lmax + 1);
    const size_t mb = sizeof(double) * npix * nmaps, ab = 2 * sizeof(double) * nalm * nmaps;
    XferBuf x(c, xfer_bytes({mb, ab}));
    double* dm = x.take(mb);
    double* da = x.take(ab);
    h2d_matrix(c, da, 2 * nalm, alms, 2 * nalm, 2 * nalm, nmaps);
    alm2map_dev(c, nside, lmax, nmaps, da, dm);
    d2h_matrix(c, maps, npix, dm, npix, npix, nmaps);
  });
}

int psx_mask_spectra(psx_ctx* h, int nside, int lmax, int niter, int nmaps, const double* maps,
                     double* cls) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  return guarded(c, [&] {
    PSX_REQUIRE(maps && cls && niter >= 0, PSX_ERR_ARGUMENT, "bad arguments");
    PSX_REQUIRE(nmaps >= 1 && nmaps <= kMaxSpectraMaps, PSX_ERR_ARGUMENT,
                "mask spectra of 1..8 maps per call");
    check_sht_args(nside, lmax);
    PSX_HIP(hipSetDevice(c->device));
    const int64_t npix = 12LL * nside * nside;
    const int64_t nalm = alm_idx(lmax, lmax + 1, lmax + 1);
    const int64_t ncl = (int64_t)nmaps * nmaps * (lmax + 1);
    const size_t mb = sizeof(double) * npix * nmaps, ab = 2 * sizeof(double) * nalm * nmaps,
                 cb = sizeof(double) * ncl;
    XferBuf x(c, xfer_bytes({mb, ab, cb}));
    double* dm = x.take(mb);
    double* da = x.take(ab);
    double* dc = x.take(cb);
    // the first analysis launch's maps go in first; the others are copied on
    // the copy stream while that launch and its ring transforms run
    const int first = std::min(nmaps, PSX_AN_GROUP);
    h2d_matrix(c, dm, npix, maps, npix, npix, first);
    const std::function<void(int)> rest = [&](int q) {
      if (q == first)
        h2d_matrix_x(c, dm + first * npix, npix, maps + first * npix, npix, npix,
                     nmaps - first);
    };
    map2alm_dev(c, nside, lmax, niter, nmaps, dm, da, first < nmaps ? &rest : nullptr);
    cross_spectra_dev(c, lmax, nmaps, da, dc);
    d2h_matrix(c, cls, ncl, dc, ncl, ncl, 1);
  });
}

int psx_alm2cl(psx_ctx* h, int lmax, int nsets, const double* const* alms, double* cls) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  return guarded(c, [&] {
    PSX_REQUIRE(alms && cls, PSX_ERR_ARGUMENT, "bad arguments");
    PSX_REQUIRE(nsets >= 1 && nsets <= kMaxSpectraMaps, PSX_ERR_ARGUMENT,
                "cross-spectra of 1..8 coefficient sets per call");
    PSX_REQUIRE(lmax >= 0 && lmax <= 8191, PSX_ERR_ARGUMENT, "lmax must be 0..8191");
    for (int i = 0; i < nsets; ++i) PSX_REQUIRE(alms[i], PSX_ERR_ARGUMENT, "null alm pointer");
    PSX_HIP(hipSetDevice(c->device));
    const int64_t nalm = alm_idx(lmax, lmax + 1, lmax + 1);
    const int64_t ncl = (int64_t)nsets * nsets * (lmax + 1);
    const size_t ab = 2 * sizeof(double) * nalm * nsets, cb = sizeof(double) * ncl;
    XferBuf x(c, xfer_bytes({ab, cb}));
    double* da = x.take(ab);
    double* dc = x.take(cb);
    for (int i = 0; i < nsets; ++i)
      h2d_matrix(c, da + 2 * nalm * i, 2 * nalm, alms[i], 2 * nalm, 2 * nalm, 1);
    cross_spectra_dev(c, lmax, nsets, da, dc);
    d2h_matrix(c, cls, ncl, dc, ncl, ncl, 1);
  });
}

int psx_master_spectra(psx_ctx* h, int nside, int lmax, int niter, const double* const* maps,
                       const double* const* masks, int already_masked, double* pcl, double* W) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  return guarded(c, [&] {
    PSX_REQUIRE(maps && masks && pcl && W && niter >= 0, PSX_ERR_ARGUMENT, "bad arguments");
    for (int i = 0; i < 6; ++i) PSX_REQUIRE(maps[i], PSX_ERR_ARGUMENT, "null map pointer");
    for (int a = 0; a < 4; ++a) PSX_REQUIRE(masks[a], PSX_ERR_ARGUMENT, "null mask pointer");
    check_sht_args(nside, lmax);
    PSX_HIP(hipSetDevice(c->device));
    const int64_t npix = 12LL * nside * nside;
    const int64_t nalm = alm_idx(lmax, lmax + 1, lmax + 1);
    const int64_t L = lmax + 1;
    // distinct masks (by address), each copied and transformed once
    const double* uniq[4];
    int slot[4], nu = 0;
    for (int a = 0; a < 4; ++a) {
      int k = 0;
      while (k < nu && uniq[k] != masks[a]) ++k;
      if (k == nu) uniq[nu++] = masks[a];
      slot[a] = k;
    }
    const size_t mb = sizeof(double) * npix * 6, wb = sizeof(double) * npix * nu,
                 ab = 2 * sizeof(double) * nalm * 6, wab = 2 * sizeof(double) * nalm * nu,
                 cb = sizeof(double) * 36 * L, wcb = sizeof(double) * nu * nu * L;
    XferBuf x(c, xfer_bytes({mb, wb, ab, wab, cb, wcb}));
    double* dm = x.take(mb);
    double* dw = x.take(wb);
    double* da = x.take(ab);
    double* dwa = x.take(wab);
    double* dc = x.take(cb);
    double* dwc = x.take(wcb);
    // map m's I by its T mask, Q and U by its P mask
    auto mask_map = [&](int i) {
      if (already_masked) return;
      const int a = 2 * (i / 3) + (i % 3 == 0 ? 0 : 1);
      hipLaunchKernelGGL(apply_mask, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0,
                         c->stream, dm + i * npix, (const double*)(dw + slot[a] * npix), npix);
      PSX_HIP(hipGetLastError());
    };
    // the I maps and the masks first: the spin-0 transforms need nothing else
    for (int i : {0, 3}) h2d_matrix(c, dm + i * npix, npix, maps[i], npix, npix, 1);
    for (int k = 0; k < nu; ++k) h2d_matrix(c, dw + k * npix, npix, uniq[k], npix, npix, 1);
    for (int i : {0, 3}) mask_map(i);
    // spin 0: the two I maps and the distinct masks, in transform groups
    {
      const double* m0[6];
      double* a0[6];
      int n0 = 0;
      for (int i = 0; i < 2; ++i, ++n0) {
        m0[n0] = dm + 3 * i * npix;
        a0[n0] = da + 6 * i * nalm;
      }
      for (int k = 0; k < nu; ++k, ++n0) {
        m0[n0] = dw + k * npix;
        a0[n0] = dwa + 2 * k * nalm;
      }
      map2alm_ptrs(c, nside, lmax, niter, n0, m0, a0, nullptr, false);
    }
    // Q and U on the copy stream while the spin-0 transforms run
    for (int i : {1, 2, 4, 5}) h2d_matrix_x(c, dm + i * npix, npix, maps[i], npix, npix, 1);
    for (int i : {1, 2, 4, 5}) mask_map(i);
    map2alm_pol_dev(c, nside, lmax, niter, 2, dm, da, false);
    cross_spectra_dev(c, lmax, 6, da, dc);
    cross_spectra_dev(c, lmax, nu, dwa, dwc);
    std::vector<double> hc(36 * L), hw((size_t)nu * nu * L);
    d2h_matrix(c, hc.data(), 36 * L, dc, 36 * L, 36 * L, 1);
    d2h_matrix(c, hw.data(), (int64_t)nu * nu * L, dwc, (int64_t)nu * nu * L,
               (int64_t)nu * nu * L, 1);
    // map 1's X with map 2's Y; mask a of map 1 with mask b of map 2
    for (int xx = 0; xx < 3; ++xx)
      for (int yy = 0; yy < 3; ++yy)
        std::copy_n(hc.data() + (xx * 6 + 3 + yy) * L, L, pcl + (xx * 3 + yy) * L);
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        std::copy_n(hw.data() + (slot[a] * nu + slot[2 + b]) * L, L, W + (a * 2 + b) * L);
  });
}

int psx_map2alm_pol(psx_ctx* h, int nside, int lmax, int niter, int nmaps, const double* iqu,
                    double* teb) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  return guarded(c, [&] {
    PSX_REQUIRE(iqu && teb && nmaps >= 1 && niter >= 0, PSX_ERR_ARGUMENT, "bad arguments");
    check_sht_args(nside, lmax);
    PSX_HIP(hipSetDevice(c->device));
    const int64_t npix = 12LL * nside * nside;
    const int64_t nalm = alm_idx(lmax, lmax + 1, lmax + 1);
    const size_t mb = 3 * sizeof(double) * npix * nmaps, ab = 6 * sizeof(double) * nalm * nmaps;
    XferBuf x(c, xfer_bytes({mb, ab}));
    double* dm = x.take(mb);
    double* da = x.take(ab);
    // the I maps first; Q and U are copied on the copy stream while I's
    // transforms run
    for (int i = 0; i < nmaps; ++i)
      h2d_matrix(c, dm + 3 * i * npix, npix, iqu + 3 * i * npix, npix, npix, 1);
    const std::function<void()> qu = [&] {
      for (int i = 0; i < nmaps; ++i)
        h2d_matrix_x(c, dm + (3 * i + 1) * npix, npix, iqu + (3 * i + 1) * npix, npix, npix, 2);
    };
    map2alm_pol_dev(c, nside, lmax, niter, nmaps, dm, da, true, &qu);
    d2h_matrix(c, teb, 2 * nalm, da, 2 * nalm, 2 * nalm, 3 * nmaps);
  });
}

int psx_alm2map_pol(psx_ctx* h, int nside, int lmax, int nmaps, const double* teb, double* iqu) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  return guarded(c, [&] {
    PSX_REQUIRE(iqu && teb && nmaps >= 1, PSX_ERR_ARGUMENT, "bad arguments");
    check_sht_args(nside, lmax);
    PSX_HIP(hipSetDevice(c->device));
    const int64_t npix = 12LL * nside * nside;
    const int64_t nalm = alm_idx(lmax, lmax + 1, lmax + 1);
    const size_t mb = 3 * sizeof(double) * npix * nmaps, ab = 6 * sizeof(double) * nalm * nmaps;
    XferBuf x(c, xfer_bytes({mb, ab}));
    double* dm = x.take(mb);
    double* da = x.take(ab);
    h2d_matrix(c, da, 2 * nalm, teb, 2 * nalm, 2 * nalm, 3 * nmaps);
    alm2map_pol_dev(c, nside, lmax, nmaps, da, dm);
    d2h_matrix(c, iqu, npix, dm, npix, npix, 3 * nmaps);
  });
}

}  // extern "C"
