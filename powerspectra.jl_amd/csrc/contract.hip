// Closed-form 3j contraction engine for gfx950: mode-coupling matrices and
// coupled covariance blocks.
//
// Replaces the per-(l1,l2) family recurrence + Ξ dot product of the reference
// (src/modecoupling.jl:3-159, src/covariance.jl:92-446).  Instead of running
// the Schulten-Gordon recurrence per pair, every 3j product the projectors
// need is written as a product of one-dimensional tables of the central
// binomial ratio gamma_m = C(2m,m)/4^m (exact identities, checked against
// sympy in tests/):
//
//   even J = l1+l2+l3 = 2g,  a = g-l1, b = g-l2, c = g-l3:
//     (l3 l1 l2; 0 0 0)^2   = gamma_a gamma_b gamma_c / ((2g+1) gamma_g)
//     (l3 l1 l2; 0 -2 2)    = (l3 l1 l2; 0 0 0) Q / (2 sqrt(Pi))
//        Q  = D0 (D0 - 2) - 2 P1 P2,  D0 = P1 + P2 - l3(l3+1),  P = l(l+1)
//        Pi = (l1-1)l1(l1+1)(l1+2) (l2-1)l2(l2+1)(l2+2)
//   odd J = 2g+1:
//     (l3 l1 l2; 0 -2 2)^2  = D1^2 z_a z_b z_c / (4 Pi gamma_{g+1})
//        z_m = (2m+1) gamma_m,  D1 = P1 + P2 - l3(l3+1) - 2
//
// With S = l1 + l2 and the summation index c (l3 = S - 2c even, S - 2c - 1
// odd; a = l2 - c, b = l1 - c, g = S - c) every term is a product of
// one-dimensional tables:
//   even:  gamma_{l1-c} * gamma_c  *  gamma_{l2-c} * delta_{S-c}  *  What_{l3}
//   odd:   z_{l1-c-1}   * z_c      *  z_{l2-c-1}   * igamma_{S-c} *  What_{l3}
// with delta_g = 1/((2g+1) gamma_g), igamma_g = 1/gamma_g and
// What_l = (2l+1) W_l / (4 pi).  The main kernel spends 2 FP64 VALU ops per
// (0 0 0)^2 term (mul + fma) -- no sqrt, no division, no recurrence.  The
// integer polynomials P, D0, D1 and kappa = 1 + 2 P1 P2 are exact doubles for
// lmax <= 8191 (2 P1 P2 <= 2 (8191 * 8192)^2 < 2^53; at 8192 it is not), and
// Q = E^2 - kappa and D1^2 take one rounding each; lmax is capped there.
//
// Work decomposition: one wavefront owns a tile of 64 consecutive l1 (lanes)
// times R consecutive anti-diagonals S = S0 .. S0+R-1 (register slots); for a
// fixed S the lanes walk along an anti-diagonal of the (l1, l2) triangle.
// Per step c:
//   * the first factor pair depends on (lane, c) only: one vector load per
//     step, times the uniform gamma_c;
//   * the second pair depends on (lane, S - c), i.e. on r - c: it lives in a
//     register window that slides by one slot per step (one vector load and
//     one multiply per step, shared by all R slots);
//   * What_{S-2c} (and l3(l3+1) for the spin-2 polynomials) is uniform across
//     the wave and comes from small 1D tables through scalar loads.
// Lane tables are read with vector memory loads (in-order vmcnt), uniform
// ones with scalar loads, so neither waits on the other.
#include "psx_internal.h"
#include "cov_epi.h"

#include <algorithm>
#include <cmath>

namespace psx {

namespace {

constexpr int kWOff = 192;   // What index of l3 = 0
constexpr int kLmaxCap = 8191;  // exactness bound of kappa, see above
constexpr double kInv4Pi = 0.07957747154594766788444188168625718;  // 1/(4 pi)

// A tile: lanes l1 = x0 .. x0+63 on anti-diagonals S0 .. S0+R-1; the tile
// owns (computes over c and writes) l1 in [lo, hi] only.
struct Tile {
  int x0, S0, lo, hi;
};

constexpr int kMaxV = 12;

struct KArgs {
  const Tile* tiles;
  int ntiles;
  int lmin, lmax;
  int LW;
  // 1D tables, index m at [m + kTabOff]
  const double* G;   // gamma_m
  const double* Z;   // (2m+1) gamma_m
  const double* D;   // 1 / ((2m+1) gamma_m)
  const double* IG;  // 1 / gamma_m
  // per-l3 records (record l3 at [(l3 + kWOff) * RS]), fields as in
  // contract_kernel
  const double* rec;
  double* out[kMaxV];
  int64_t ldo;
  const double* spec[4];
  const double* ratio[4];
  int* err;
  const int64_t* doff;   // EPI_RMAX
};

template <typename T>
__device__ __forceinline__ T sload(const T* p, int64_t i) {
  // Uniform address: force a scalar (SMEM) load through the constant
  // address space so the value lands in SGPRs.
  return ((const __attribute__((address_space(4))) T*)p)[i];
}

__device__ __forceinline__ int ufl(int v) { return __builtin_amdgcn_readfirstlane(v); }

// ---- the contraction kernel ---------------------------------------------
// Uniform per-l3 data lives in one interleaved record table: record l3 holds
// RS doubles, fields
//   [0, NE)         What_k(l3)       for the even-family tables (00, 02, ++)
//   [NE, NV)        What_k(l3 - 1)   for the odd-family tables (--)
//   FT  = NV        l3 (l3 + 1)                      (if a 02/++ table is present)
//   FT1 = NV + kE   (l3 - 1) l3 + 1                  (if a -- table is present)
// so the R records a step needs are R consecutive records behind one base
// pointer: few scalar loads, one address register, immediate offsets.
__host__ __device__ constexpr int rec_fields(int nv, bool e, bool z) {
  return nv + (e ? 1 : 0) + (z ? 1 : 0);
}
__host__ __device__ constexpr int rec_stride(int nf) {
  return nf <= 1 ? 1 : nf <= 2 ? 2 : nf <= 4 ? 4 : nf <= 8 ? 8 : 16;
}

// One contracted term: acc += w t, or for EPI_RMAX the running max of |w t|.
template <int EPI>
__device__ __forceinline__ double accum(double w, double t, double acc) {
  if constexpr (EPI == EPI_RMAX) return fmax(acc, fabs(w * t));
  else return fma(w, t, acc);
}

// EPI: EPI_MCM, EPI_COV or EPI_RMAX; COVK: covariance kind for EPI_COV; WPE: minimum
// waves per SIMD the register allocation must allow; TPB: threads per
// workgroup (64 per tile).
template <int R, int N00, int N02, int NPE, int NPO, int EPI, int COVK, int WPE, int TPB>
__global__ void __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(WPE)))
contract_kernel(KArgs A) {
  constexpr int NV = N00 + N02 + NPE + NPO;
  constexpr int NE = N00 + N02 + NPE;
  constexpr bool kG = NE > 0;
  constexpr bool kZ = NPO > 0;
  constexpr bool kE = (N02 + NPE) > 0;
  constexpr bool kS = kE || kZ;  // spin-2 polynomials needed
  constexpr int RS = rec_stride(rec_fields(NV, kE, kZ));
  constexpr int FT = NV, FT1 = NV + (kE ? 1 : 0);

  const int lane = threadIdx.x & 63;
  const int wave = ufl(threadIdx.x >> 6);
  // tile lists come in whole workgroups (padded with empty tiles): every wave
  // of a launched workgroup has a tile and reaches the epilogue's barriers
  const int tile = ufl(blockIdx.x * (TPB / 64) + wave);
  if (ufl(blockIdx.x * (TPB / 64)) >= A.ntiles) return;
  const int x0 = sload(&A.tiles[tile].x0, 0);
  const int S0 = sload(&A.tiles[tile].S0, 0);
  const int own_lo = sload(&A.tiles[tile].lo, 0);
  const int own_hi = sload(&A.tiles[tile].hi, 0);
  const int lmax = A.lmax;
  const int l1 = x0 + lane;

  // summation range over c (c <= min(l1, l2) = l1 <= own_hi), rounded out
  // to whole blocks of R
  const int chi = own_hi;
  const int clo = max(0, (S0 - A.LW + 1) >> 1);
  const int cbeg = ufl((clo / R) * R);
  const int nblk = ufl(max(0, (chi + 1 - cbeg + R - 1) / R));

  // lane constants per slot: E = sig[r] - T(l3), Q = E^2 - kap[r],
  // D1 = sig[r] - T1(l3 - 1)
  double sig[kS ? R : 1], kap[kE ? R : 1];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const double L1 = (double)l1, L2 = (double)(S0 + r - l1);
    const double P1 = L1 * (L1 + 1.0), P2 = L2 * (L2 + 1.0);
    if constexpr (kS) sig[r] = P1 + P2 - 1.0;
    if constexpr (kE) kap[r] = 1.0 + 2.0 * P1 * P2;
  }

  double acc[NV][R];
#pragma unroll
  for (int k = 0; k < NV; ++k)
#pragma unroll
    for (int r = 0; r < R; ++r) acc[k][r] = 0.0;

  // Walking pointers; step c = cbeg + blk R + u reads
  //   lane:    G[l1 - c] = gl[-u],  Z[l1 - c - 1] = zl[-u]
  //   uniform: gamma_c = sg[u],  z_c = sz[u]
  //   window refill (q = -c - 1): G[t + q] = gwp[-u], Z[t + q - 1] = zwp[-u],
  //            delta_{S0+q} = sdp[-u], igamma_{S0+q} = sip[-u]
  //   records: l3 = S0 + r - 2c at rec[(r - 2u) RS]
  // with t = S0 - x0 - lane; every pointer advances once per block.
  const int t = S0 - x0 - lane;
  const double* gl = A.G + kTabOff + (l1 - cbeg);
  const double* zl = A.Z + kTabOff + (l1 - cbeg - 1);
  const double* gwp = A.G + kTabOff + (t - cbeg - 1);
  const double* zwp = A.Z + kTabOff + (t - cbeg - 2);
  const double* sg = A.G + kTabOff + cbeg;
  const double* sz = A.Z + kTabOff + cbeg;
  const double* sdp = A.D + kTabOff + (S0 - cbeg - 1);
  const double* sip = A.IG + kTabOff + (S0 - cbeg - 1);
  const double* rec = A.rec + (int64_t)(kWOff + S0 - 2 * cbeg) * RS;

  // sliding windows: slot r at step c needs (q = r - c)
  //   gw: gamma_{t+q} delta_{S0+q},  zw: z_{t+q-1} igamma_{S0+q},
  // held at index q mod R (cbeg % R == 0).
  double gw[kG ? R : 1], zw[kZ ? R : 1];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if constexpr (kG) gw[r] = gwp[r + 1] * sload(sdp, r + 1);
    if constexpr (kZ) zw[r] = zwp[r + 1] * sload(sip, r + 1);
  }

  for (int blk = 0; blk < nblk; ++blk) {
#pragma unroll
    for (int u = 0; u < R; ++u) {
      double g1 = 0.0, z1 = 0.0;
      if constexpr (kG) g1 = gl[-u] * sload(sg, u);
      if constexpr (kZ) z1 = zl[-u] * sload(sz, u);
      const double* ru = rec - 2 * u * RS;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int wi = ((r - u) % R + R) % R;
        const double* rr = ru + r * RS;
        double p = 0.0, Q = 0.0;
        if constexpr (kG) p = g1 * gw[wi];
        if constexpr (kE) {
          const double E = sig[r] - sload(rr, FT);
          Q = fma(E, E, -kap[r]);
        }
#pragma unroll
        for (int k = 0; k < N00; ++k) acc[k][r] = accum<EPI>(sload(rr, k), p, acc[k][r]);
        double pq = 0.0;
        if constexpr (N02 > 0) {
          pq = p * Q;
#pragma unroll
          for (int k = 0; k < N02; ++k)
            acc[N00 + k][r] = accum<EPI>(sload(rr, N00 + k), pq, acc[N00 + k][r]);
        }
        if constexpr (NPE > 0) {
          const double pqq = (N02 > 0 ? pq : p * Q) * Q;
#pragma unroll
          for (int k = 0; k < NPE; ++k)
            acc[N00 + N02 + k][r] =
                accum<EPI>(sload(rr, N00 + N02 + k), pqq, acc[N00 + N02 + k][r]);
        }
        if constexpr (NPO > 0) {
          const double D1 = sig[r] - sload(rr, FT1);
          const double pd = (z1 * zw[wi]) * (D1 * D1);
#pragma unroll
          for (int k = 0; k < NPO; ++k)
            acc[NE + k][r] = accum<EPI>(sload(rr, NE + k), pd, acc[NE + k][r]);
        }
      }
      // advance the windows to step c + 1: the slot that held r = R-1 now
      // receives q = -c - 1
      const int ni = (R - 1 - u) % R;
      if constexpr (kG) gw[ni] = gwp[-u] * sload(sdp, -u);
      if constexpr (kZ) zw[ni] = zwp[-u] * sload(sip, -u);
    }
    gl -= R;
    zl -= R;
    gwp -= R;
    zwp -= R;
    sg += R;
    sz += R;
    sdp -= R;
    sip -= R;
    rec -= 2 * R * RS;
  }

  // ---- epilogue ----
  if constexpr (EPI == EPI_RMAX) {
    // the row maxima, one per owned pair, anti-diagonal-major
    const int lmin = A.lmin;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int S = S0 + r, l2 = S - l1;
      if (l1 < own_lo || l1 > own_hi || l1 < lmin || l1 > l2 || l2 > lmax) continue;
      A.out[0][A.doff[S - 2 * lmin] + (l1 - max(lmin, S - lmax))] = acc[0][r];
    }
    return;
  }
  // The workgroup's tiles are one row chunk over WG R consecutive
  // anti-diagonals: S0 of wave w = S0 of wave 0 + w R, empty tiles (hi < lo)
  // last (checked on the host).  Each output matrix goes through LDS and
  // leaves as column segments: column l1 takes rows l2 = S - l1 (the mirrored
  // element) and column l2 rows l1 = S - l2 (the direct one) over the span of
  // anti-diagonals, consecutive threads on consecutive rows, instead of one
  // scattered 8-byte store per lane.
  constexpr int WG = TPB / 64, SPAN = WG * R, CPP = TPB / SPAN;
  static_assert(TPB % SPAN == 0, "anti-diagonal span must divide the workgroup");
  __shared__ double xs[WG][R][64];
  __shared__ int tinfo[WG][4];
  const int lmin = A.lmin;
  if (lane == 0) {
    tinfo[wave][0] = x0;
    tinfo[wave][1] = S0;
    tinfo[wave][2] = own_lo;
    tinfo[wave][3] = own_hi;
  }
  const double L1 = (double)l1;
  const double pi1 = (L1 - 1.0) * L1 * (L1 + 1.0) * (L1 + 2.0);
  double k02r[kE ? R : 1], k22r[(NPE + NPO) > 0 ? R : 1];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int l2 = S0 + r - l1;
    const double L2 = (double)l2;
    const double pi2 = (L2 - 1.0) * L2 * (L2 + 1.0) * (L2 + 2.0);
    const bool spin2 = (l1 >= 2) && (l2 >= 2);
    if constexpr (kE) k02r[r] = spin2 ? 1.0 / (2.0 * sqrt(pi1 * pi2)) : 0.0;
    if constexpr ((NPE + NPO) > 0) k22r[r] = spin2 ? 1.0 / (4.0 * pi1 * pi2) : 0.0;
  }
  bool bad = false;
  // the value of pair (l1, S - l1) if this workgroup computed and owns it
  auto staged = [&](int S, int m1, double& v) -> bool {
    const int i = S - tinfo[0][1];
    if (i < 0 || i >= SPAN) return false;
    const int w = i / R, r = i - w * R;
    if (m1 < tinfo[w][2] || m1 > tinfo[w][3]) return false;
    const int m2 = S - m1;
    if (m1 < lmin || m1 > m2 || m2 > lmax) return false;
    v = xs[w][r][m1 - tinfo[w][0]];
    return true;
  };
  // column segments of one output matrix; MCM weights (2 l2 + 1) on the
  // direct element, (2 l1 + 1) on the mirrored one, none for covariance
  auto store = [&](double* o, bool weighted) {
    __syncthreads();
    int l1lo = lmax + 1, l1hi = -1;
#pragma unroll
    for (int w = 0; w < WG; ++w)
      if (tinfo[w][3] >= tinfo[w][2]) {
        l1lo = min(l1lo, tinfo[w][2]);
        l1hi = max(l1hi, tinfo[w][3]);
      }
    const int s0 = tinfo[0][1];
    const int i = threadIdx.x % SPAN, cs = threadIdx.x / SPAN;
    const int S = s0 + i;
    for (int m1 = l1lo + cs; m1 <= l1hi; m1 += CPP) {
      double v;
      if (staged(S, m1, v))
        o[(int64_t)(S - m1 - lmin) + (int64_t)(m1 - lmin) * A.ldo] =
            weighted ? (2.0 * (double)m1 + 1.0) * v : v;
    }
    for (int m2 = s0 - l1hi + cs; m2 <= s0 + SPAN - 1 - l1lo; m2 += CPP) {
      double v;
      if (staged(S, S - m2, v))
        o[(int64_t)(S - m2 - lmin) + (int64_t)(m2 - lmin) * A.ldo] =
            weighted ? (2.0 * (double)m2 + 1.0) * v : v;
    }
    __syncthreads();
  };
  if constexpr (EPI == EPI_MCM) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        double kap_k = 1.0;
        if constexpr (N02 > 0)
          if (k >= N00 && k < N00 + N02) kap_k = k02r[r];
        if constexpr ((NPE + NPO) > 0)
          if (k >= N00 + N02) kap_k = k22r[r];
        xs[wave][r][lane] = kap_k * acc[k][r];
      }
      store(A.out[k], true);
    }
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int l2 = S0 + r - l1;
      double v = 0.0;
      if (!(l1 < own_lo || l1 > own_hi || l1 < lmin || l1 > l2 || l2 > lmax)) {
        double xi[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) {
          double kap_k = 1.0;
          if constexpr (N02 > 0)
            if (k >= N00 && k < N00 + N02) kap_k = k02r[r];
          if constexpr ((NPE + NPO) > 0)
            if (k >= N00 + N02) kap_k = k22r[r];
          xi[k] = kap_k * acc[k][r];
        }
        v = cov_combine<COVK>(l1, l2, xi, A.spec, A.ratio, bad);
      }
      xs[wave][r][lane] = v;
    }
    store(A.out[0], false);
    if (bad) atomicOr(A.err, 1);
  }
}

// ---- record table: What_k = (2 l3 + 1) W_k[l3] / (4 pi), zero outside
// [0, LW), and the l3 polynomials, interleaved as described above ----------
struct RecArgs {
  double* rec;
  const double* W[kMaxV];
  int len;  // records (index l3 + kWOff)
  int LW;
  int nv, ne, rs;
  int ft, ft1;  // field of T / T1, or -1
  int* err;     // cleared here (covariance calls), or null
  int raw;      // V = W itself (EPI_RMAX), not (2 l3 + 1) W / (4 pi)
};

__device__ __forceinline__ double what_at(const double* W, int l3, int LW, int raw) {
  if (l3 < 0 || l3 >= LW) return 0.0;
  return raw ? W[l3] : (2.0 * l3 + 1.0) * W[l3] * kInv4Pi;
}

__global__ void __launch_bounds__(256) record_kernel(RecArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0 && a.err) *a.err = 0;
  if (i >= a.len) return;
  const int l3 = i - kWOff;
  double* out = a.rec + (int64_t)i * a.rs;
  for (int f = 0; f < a.rs; ++f) {
    double v = 0.0;
    if (f < a.ne) {
      v = what_at(a.W[f], l3, a.LW, a.raw);
    } else if (f < a.nv) {
      v = what_at(a.W[f], l3 - 1, a.LW, a.raw);
    } else if (f == a.ft) {
      v = (double)l3 * (double)(l3 + 1);
    } else if (f == a.ft1) {
      v = (double)(l3 - 1) * (double)l3 + 1.0;
    }
    out[f] = v;
  }
}

// ---- tile lists ----------------------------------------------------------
struct TileKey {
  int lmin, lmax, LW, R, waves, nbands;
  bool operator<(const TileKey& o) const {
    return std::tie(lmin, lmax, LW, R, waves, nbands) <
           std::tie(o.lmin, o.lmax, o.LW, o.R, o.waves, o.nbands);
  }
};

// Tiles follow the anti-diagonals: for each group S0 .. S0+R-1 the owned
// range l1 in [max(lmin, S0-lmax), min((S0+R-1)/2, lmax)] (the l1 <= l2
// half of the group) is cut into chunks of 64 from the diagonal down, so
// only the cheap chunk nearest l1 = lmin (or l2 = lmax) is partial.  Lanes
// never go below l1 = 0: a bottom chunk that would is shifted up and owns
// only its part.
//
// Launch order: the `waves` tiles of one workgroup are the same chunk (k-th
// from the diagonal) of `waves` consecutive groups, so one CU writes
// waves*R consecutive anti-diagonals of the same l1 range and the output
// column segments it produces (waves*R doubles) fill whole cache lines in
// one L2.  Workgroups are sorted by cost (c-range length), longest first.
std::vector<Tile> make_tiles(int lmin, int lmax, int LW, int R, int waves) {
  struct TC {
    Tile t;
    int cost;
  };
  // chunks[k] = tiles k-th from the diagonal, in S0 order
  std::vector<std::vector<TC>> chunks;
  for (int S0 = 2 * lmin; S0 <= 2 * lmax; S0 += R) {
    const int hi = std::min((S0 + R - 1) / 2, lmax);
    const int lo = std::max(lmin, S0 - lmax);
    int k = 0;
    for (int top = hi; top >= lo; top -= 64, ++k) {
      Tile t;
      t.x0 = std::max(top - 63, 0);
      t.S0 = S0;
      t.lo = std::max(lo, top - 63);
      t.hi = top;
      const int clo = std::max(0, (S0 - LW + 1) >> 1);
      const int cbeg = (clo / R) * R;
      const int nblk = std::max(0, (t.hi + 1 - cbeg + R - 1) / R);
      if ((int)chunks.size() <= k) chunks.resize(k + 1);
      chunks[k].push_back({t, nblk});
    }
  }
  struct WG {
    int first, n, cost;
  };
  std::vector<TC> flat;
  std::vector<WG> wgs;
  for (auto& ck : chunks) {
    for (size_t i = 0; i < ck.size(); i += waves) {
      WG g{(int)flat.size(), 0, 0};
      for (size_t j = i; j < std::min(ck.size(), i + waves); ++j) {
        flat.push_back(ck[j]);
        g.n++;
        g.cost = std::max(g.cost, ck[j].cost);
      }
      wgs.push_back(g);
    }
  }
  // Longest first, in coarse cost buckets; inside a bucket by (chunk, S0) so
  // neighbouring anti-diagonal groups stay neighbours in the order.
  int maxc = 1;
  for (const WG& g : wgs) maxc = std::max(maxc, g.cost);
  const int bucket = std::max(1, maxc / 64);
  std::stable_sort(wgs.begin(), wgs.end(), [bucket](const WG& a, const WG& b) {
    return a.cost / bucket > b.cost / bucket;
  });
  // Workgroup b runs on XCD b mod kXcds.  Give each XCD runs of kRun
  // consecutive workgroups of the order above, so the partial output lines
  // two neighbouring groups write meet in the same L2; XCDs still take
  // turns along the cost order.
  constexpr int kXcds = 8, kRun = 8;
  const int nwg = (int)wgs.size();
  std::vector<int> perm;
  perm.reserve(nwg);
  for (int b = 0; (int)perm.size() < nwg; ++b) {
    const int x = b % kXcds, m = b / kXcds;
    const long w = (long)(m / kRun) * kXcds * kRun + (long)x * kRun + m % kRun;
    if (w < nwg) perm.push_back((int)w);
    if (b > 4 * nwg + 64 * kXcds * kRun) break;  // unreachable guard
  }
  // every workgroup occupies exactly `waves` slots; short groups are padded
  // with empty tiles (hi = -1 < lo: no lane owns anything, no c step runs)
  std::vector<Tile> out;
  out.reserve((size_t)nwg * waves);
  for (int w : perm) {
    const WG& g = wgs[w];
    for (int j = 0; j < waves; ++j) {
      if (j < g.n) {
        out.push_back(flat[g.first + j].t);
      } else {
        out.push_back(Tile{0, 2 * lmin, 0, -1});
      }
    }
  }
  return out;
}

// One l1 band of a tile list: workgroups [wg0, wg0 + nwg).  Once this band
// and every earlier one have run, every pair with l1 < done_l is written, and
// with it every output column l < done_l: column l holds pairs (l1, l) with
// l1 <= l and, mirrored, pairs (l, l2) with l2 > l, so it needs exactly the
// pairs with l1 <= l.
struct TileBand {
  int wg0, nwg, done_l;
};

// Reorder the workgroups of `tiles` (blocks of `waves` tiles) into nbands l1
// bands by the top lane of their highest tile, keeping the launch order
// inside a band.  Band edges sit at 1/8, 1/4, 1/2 and 3/4 of the l range: the
// work per l1 row grows like l1 (lmax - l1), so the first band is cheap and
// the columns it completes can be copied out while the rest runs.
std::vector<TileBand> band_tiles(std::vector<Tile>& tiles, int waves, int lmin, int lmax,
                                 int nbands) {
  const int nwg = (int)(tiles.size() / waves);
  if (nbands <= 1) return {TileBand{0, nwg, lmax + 1}};
  static const double kEdge[4] = {0.125, 0.25, 0.5, 0.75};
  const int n = lmax - lmin + 1;
  std::vector<int> edge;
  for (int k = 0; k < nbands - 1 && k < 4; ++k) edge.push_back(lmin + (int)(kEdge[k] * n));
  const int nb = (int)edge.size() + 1;
  auto band_of = [&](int w) {
    int top = -1;
    for (int j = 0; j < waves; ++j) top = std::max(top, tiles[(size_t)w * waves + j].hi);
    int b = 0;
    while (b < nb - 1 && top >= edge[b]) ++b;
    return b;
  };
  std::vector<std::vector<int>> members(nb);
  for (int w = 0; w < nwg; ++w) members[band_of(w)].push_back(w);
  std::vector<Tile> out;
  out.reserve(tiles.size());
  std::vector<TileBand> bands;
  for (int b = 0; b < nb; ++b) {
    TileBand tb{(int)(out.size() / waves), (int)members[b].size(), lmax + 1};
    for (int w : members[b])
      for (int j = 0; j < waves; ++j) out.push_back(tiles[(size_t)w * waves + j]);
    bands.push_back(tb);
  }
  // done_l of band b: the lowest l1 any later band still owns
  int low = lmax + 1;
  for (int b = nb - 1; b >= 0; --b) {
    bands[b].done_l = low;
    const TileBand& tb = bands[b];
    for (size_t i = (size_t)tb.wg0 * waves; i < (size_t)(tb.wg0 + tb.nwg) * waves; ++i)
      if (out[i].hi >= out[i].lo) low = std::min(low, out[i].lo);
  }
  tiles.swap(out);
  return bands;
}

struct Geometry {
  int R;
  int wh_len;     // records (l3 = -kWOff .. wh_len - kWOff - 1)
  int tab_need;   // largest 1D table index any read touches
};

// Index ranges touched by contract_kernel for this tile set, checked against
// the buffers before launch (the access list follows the kernel: lane reads
// of G/Z at l1-c(-1) and t+q(-1), uniform reads of D/IG at S0+q, of G/Z at c,
// of the l3 records at l3 = S0+r-2c).
Geometry check_geometry(const std::vector<Tile>& tiles, int lmax, int LW, int R, int waves,
                        int tables_n) {
  Geometry g{};
  g.R = R;
  // workgroup invariants of the epilogue: whole workgroups, non-empty tiles
  // first, wave j on anti-diagonals S0 + j R of its first tile
  PSX_REQUIRE(waves >= 1 && tiles.size() % waves == 0, PSX_ERR_ARGUMENT,
              "tile geometry: not whole workgroups");
  for (size_t w0 = 0; w0 < tiles.size(); w0 += waves) {
    bool empty_seen = false;
    for (int j = 0; j < waves; ++j) {
      const Tile& t = tiles[w0 + j];
      if (t.hi < t.lo) {
        empty_seen = true;
        continue;
      }
      PSX_REQUIRE(!empty_seen && t.S0 == tiles[w0].S0 + j * R, PSX_ERR_ARGUMENT,
                  "tile geometry: workgroup anti-diagonals not consecutive");
    }
  }
  long lo = 0, hi = 0, l3lo = 0, l3hi = 0;
  bool first = true;
  for (const Tile& t : tiles) {
    PSX_REQUIRE(t.x0 >= 0 && t.lo >= t.x0 && t.hi <= t.x0 + 63 && t.hi <= lmax,
                PSX_ERR_ARGUMENT, "tile geometry: bad lane range");
    const int chi = t.hi;
    const int clo = std::max(0, (t.S0 - LW + 1) >> 1);
    const int cbeg = (clo / R) * R;
    const int nblk = std::max(0, (chi + 1 - cbeg + R - 1) / R);
    const int cend = cbeg + nblk * R;  // exclusive
    const long tmin = (long)t.S0 - t.x0 - 63, tmax = (long)t.S0 - t.x0;
    const long a = std::min({(long)t.x0 - cend, tmin - cend - 1, tmin - cbeg - 1,
                             (long)t.S0 - cend, (long)cbeg});
    const long b = std::max({(long)t.x0 + 63 - cbeg, tmax + R - 1 - cbeg,
                             (long)t.S0 + R - 1 - cbeg, (long)cend - 1});
    const long e = (long)t.S0 - 2L * (cend - 1) - 1;
    const long f = (long)t.S0 + R - 1 - 2L * cbeg;
    if (first) {
      lo = a; hi = b; l3lo = e; l3hi = f;
      first = false;
    } else {
      lo = std::min(lo, a); hi = std::max(hi, b);
      l3lo = std::min(l3lo, e); l3hi = std::max(l3hi, f);
    }
  }
  PSX_REQUIRE(lo >= -kTabOff && l3lo >= -kTabOff, PSX_ERR_ARGUMENT,
              "tile geometry: table underflow");
  PSX_REQUIRE(l3lo >= -kWOff, PSX_ERR_ARGUMENT, "tile geometry: What underflow");
  g.wh_len = (int)(((long)kWOff + std::max<long>(l3hi + 1, LW) + 15) & ~15L);
  g.tab_need = (int)std::max(hi, l3hi);
  PSX_REQUIRE(tables_n == 0 || g.tab_need + kTabOff < tables_n, PSX_ERR_ARGUMENT,
              "tile geometry: table overflow");
  return g;
}

struct CachedTiles {
  Tile* d = nullptr;
  int n = 0;
  std::vector<Tile> host;
  std::vector<TileBand> bands;  // in launch order, covering all workgroups
  Geometry geo;  // checked once per tile list (the table size is checked per call)
};

std::map<std::pair<void*, TileKey>, CachedTiles> g_tile_cache;
std::mutex g_tile_mu;

}  // namespace

// Forget the tile lists of a context that is being destroyed (their device
// copies go with the context's allocations; a later context could reuse the
// address).
void release_tile_cache(Ctx* c) {
  std::lock_guard<std::mutex> lk(g_tile_mu);
  for (auto it = g_tile_cache.begin(); it != g_tile_cache.end();) {
    if (it->first.first == (void*)c)
      it = g_tile_cache.erase(it);
    else
      ++it;
  }
}

namespace {

const CachedTiles& get_tiles(Ctx* c, int lmin, int lmax, int LW, int R, int waves,
                             int nbands) {
  std::lock_guard<std::mutex> lk(g_tile_mu);
  auto key = std::make_pair((void*)c, TileKey{lmin, lmax, LW, R, waves, nbands});
  auto it = g_tile_cache.find(key);
  if (it != g_tile_cache.end()) return it->second;
  CachedTiles ct;
  ct.host = make_tiles(lmin, lmax, LW, R, waves);
  ct.bands = band_tiles(ct.host, waves, lmin, lmax, nbands);
  ct.geo = check_geometry(ct.host, lmax, LW, R, waves, 0);
  ct.n = (int)ct.host.size();
  ct.d = (Tile*)ctx_alloc(c, sizeof(Tile) * std::max(1, ct.n));
  PSX_HIP(hipMemcpy(ct.d, ct.host.data(), sizeof(Tile) * ct.n, hipMemcpyHostToDevice));
  return g_tile_cache.emplace(key, std::move(ct)).first->second;
}

// ---- configuration per request --------------------------------------------
using KernelFn = void (*)(KArgs);

struct Config {
  int R;
  int waves;  // per workgroup
  KernelFn fn;
  const char* name;
};

template <int R, int N00, int N02, int NPE, int NPO, int EPI, int COVK, int WPE = 1,
          int WAVES = 4>
Config cfg(const char* name) {
  return Config{R, WAVES, contract_kernel<R, N00, N02, NPE, NPO, EPI, COVK, WPE, 64 * WAVES>,
                name};
}

// Anti-diagonals per tile (R) and minimum waves per SIMD for each kernel,
// chosen by timing at lmax = 3071 on gfx950.
#ifndef PSX_R_00
#define PSX_R_00 16
#endif
#ifndef PSX_R_02
#define PSX_R_02 8
#endif
#ifndef PSX_R_PP
#define PSX_R_PP 8
#endif
#ifndef PSX_R_MM
#define PSX_R_MM 8
#endif
#ifndef PSX_R_PM
#define PSX_R_PM 4
#endif
#ifndef PSX_R_SET
#define PSX_R_SET 4
#endif
#ifndef PSX_WPE_00
#define PSX_WPE_00 1
#endif
#ifndef PSX_WPE_02
#define PSX_WPE_02 1
#endif
#ifndef PSX_WPE_PM
#define PSX_WPE_PM 1
#endif
#ifndef PSX_WPE_SET
#define PSX_WPE_SET 4
#endif
#ifndef PSX_WPE_COV
#define PSX_WPE_COV 1
#endif
#ifndef PSX_R_COV
#define PSX_R_COV 4
#endif
#ifndef PSX_WAVES_SET
#define PSX_WAVES_SET 4
#endif

Config pick_config(const ContractRequest& rq) {
  const int* f = rq.nfam;
  if (rq.epi == EPI_RMAX) {
    // row maxima of one family against one raw table (the matrix-core
    // contraction's row exponents)
    if (f[0] == 1 && f[1] == 0 && f[2] == 0 && f[3] == 0)
      return cfg<8, 1, 0, 0, 0, EPI_RMAX, 0>("psx_rowmax");
    if (f[0] == 0 && f[1] == 1 && f[2] == 0 && f[3] == 0)
      return cfg<8, 0, 1, 0, 0, EPI_RMAX, 0>("psx_rowmax");
    if (f[0] == 0 && f[1] == 0 && f[2] == 1 && f[3] == 0)
      return cfg<8, 0, 0, 1, 0, EPI_RMAX, 0>("psx_rowmax");
    if (f[0] == 0 && f[1] == 0 && f[2] == 0 && f[3] == 1)
      return cfg<8, 0, 0, 0, 1, EPI_RMAX, 0>("psx_rowmax");
    throw Error{PSX_ERR_ARGUMENT, "row maxima of one family only"};
  }
  if (rq.epi == EPI_MCM) {
    if (f[0] == 1 && f[1] == 0 && f[2] == 0 && f[3] == 0) {
      // TT: 8 anti-diagonals per tile up to 4096 rows (5 % faster at lmax 3071),
      // 16 above (4.5 % faster at lmax 6143)
      if (rq.lmax - rq.lmin < 4096)
        return cfg<8, 1, 0, 0, 0, EPI_MCM, 0, PSX_WPE_00>("psx_mcm_00");
      return cfg<PSX_R_00, 1, 0, 0, 0, EPI_MCM, 0, PSX_WPE_00>("psx_mcm_00");
    }
    if (f[0] == 0 && f[1] == 1 && f[2] == 0 && f[3] == 0)
      return cfg<PSX_R_02, 0, 1, 0, 0, EPI_MCM, 0, PSX_WPE_02>("psx_mcm_02");
    if (f[0] == 0 && f[1] == 0 && f[2] == 1 && f[3] == 0)
      return cfg<PSX_R_PP, 0, 0, 1, 0, EPI_MCM, 0, PSX_WPE_02>("psx_mcm_pp");
    if (f[0] == 0 && f[1] == 0 && f[2] == 0 && f[3] == 1)
      return cfg<PSX_R_MM, 0, 0, 0, 1, EPI_MCM, 0, PSX_WPE_02>("psx_mcm_mm");
    if (f[0] == 0 && f[1] == 0 && f[2] == 1 && f[3] == 1)
      return cfg<PSX_R_PM, 0, 0, 1, 1, EPI_MCM, 0, PSX_WPE_PM>("psx_mcm_pm");
    if (f[0] == 1 && f[1] == 2 && f[2] == 1 && f[3] == 1)
      return cfg<PSX_R_SET, 1, 2, 1, 1, EPI_MCM, 0, PSX_WPE_SET, PSX_WAVES_SET>("psx_mcm_set");
    if (f[0] == 1 && f[1] == 1 && f[2] == 1 && f[3] == 1)
      return cfg<PSX_R_SET, 1, 1, 1, 1, EPI_MCM, 0, PSX_WPE_SET>("psx_mcm_set1");
    // groups of 2, 4 or 8 spectra of one family (batches of blocks in FP64)
    if (f[1] == 0 && f[2] == 0 && f[3] == 0) {
      if (f[0] == 2) return cfg<4, 2, 0, 0, 0, EPI_MCM, 0>("psx_mcm_group");
      if (f[0] == 4) return cfg<4, 4, 0, 0, 0, EPI_MCM, 0>("psx_mcm_group");
      if (f[0] == 8) return cfg<4, 8, 0, 0, 0, EPI_MCM, 0>("psx_mcm_group");
    }
    if (f[0] == 0 && f[2] == 0 && f[3] == 0) {
      if (f[1] == 2) return cfg<4, 0, 2, 0, 0, EPI_MCM, 0>("psx_mcm_group");
      if (f[1] == 4) return cfg<4, 0, 4, 0, 0, EPI_MCM, 0>("psx_mcm_group");
      if (f[1] == 8) return cfg<4, 0, 8, 0, 0, EPI_MCM, 0>("psx_mcm_group");
    }
    if (f[0] == 0 && f[1] == 0 && f[3] == 0) {
      if (f[2] == 2) return cfg<4, 0, 0, 2, 0, EPI_MCM, 0>("psx_mcm_group");
      if (f[2] == 4) return cfg<4, 0, 0, 4, 0, EPI_MCM, 0>("psx_mcm_group");
      if (f[2] == 8) return cfg<4, 0, 0, 8, 0, EPI_MCM, 0>("psx_mcm_group");
    }
    if (f[0] == 0 && f[1] == 0 && f[2] == 0) {
      if (f[3] == 2) return cfg<4, 0, 0, 0, 2, EPI_MCM, 0>("psx_mcm_group");
      if (f[3] == 4) return cfg<4, 0, 0, 0, 4, EPI_MCM, 0>("psx_mcm_group");
      if (f[3] == 8) return cfg<4, 0, 0, 0, 8, EPI_MCM, 0>("psx_mcm_group");
    }
  } else {
    constexpr int RC = PSX_R_COV;
    switch (rq.cov_kind) {
      case PSX_COV_TTTT: return cfg<RC, 8, 0, 0, 0, EPI_COV, PSX_COV_TTTT, PSX_WPE_COV>("psx_cov_tttt");
      case PSX_COV_EEEE: return cfg<RC, 0, 0, 8, 0, EPI_COV, PSX_COV_EEEE, PSX_WPE_COV>("psx_cov_eeee");
      case PSX_COV_TTTE: return cfg<RC, 4, 0, 0, 0, EPI_COV, PSX_COV_TTTE, PSX_WPE_COV>("psx_cov_ttte");
      case PSX_COV_TETE: return cfg<RC, 1, 4, 0, 0, EPI_COV, PSX_COV_TETE, PSX_WPE_COV>("psx_cov_tete");
      case PSX_COV_TEEE: return cfg<RC, 0, 4, 0, 0, EPI_COV, PSX_COV_TEEE, PSX_WPE_COV>("psx_cov_teee");
      case PSX_COV_TEEE_PLANCK:
        return cfg<RC, 0, 0, 4, 0, EPI_COV, PSX_COV_TEEE_PLANCK, PSX_WPE_COV>("psx_cov_teee_planck");
      case PSX_COV_TTEE: return cfg<RC, 2, 0, 0, 0, EPI_COV, PSX_COV_TTEE, PSX_WPE_COV>("psx_cov_ttee");
    }
  }
  throw Error{PSX_ERR_ARGUMENT, "unsupported contraction configuration"};
}

int nv_of(const ContractRequest& rq) {
  return rq.nfam[0] + rq.nfam[1] + rq.nfam[2] + rq.nfam[3];
}

}  // namespace

static int* run_contraction_checked(Ctx* c, const ContractRequest& rq, hipStream_t s,
                                    void* scratch, size_t scratch_bytes, int nbands = 1,
                                    std::vector<std::pair<int, hipEvent_t>>* band_done =
                                        nullptr);

struct RecLayout {
  int nv, ne, rs, ft, ft1;
};

RecLayout rec_layout(const ContractRequest& rq) {
  const int* f = rq.nfam;
  RecLayout L{};
  L.nv = f[0] + f[1] + f[2] + f[3];
  L.ne = f[0] + f[1] + f[2];
  const bool e = (f[1] + f[2]) > 0, z = f[3] > 0;
  L.rs = rec_stride(rec_fields(L.nv, e, z));
  L.ft = e ? L.nv : -1;
  L.ft1 = z ? L.nv + (e ? 1 : 0) : -1;
  return L;
}

size_t contraction_scratch_bytes(const ContractRequest& rq) {
  // record table (len x rs doubles) + error flag
  const int64_t LW = std::min<int64_t>(rq.LW, 2 * (int64_t)rq.lmax + 1);
  const int64_t len = kWOff + LW + 3 * 16 + 32;
  return (size_t)rec_layout(rq).rs * len * sizeof(double) + 512;
}

int* run_contraction_on(Ctx* c, const ContractRequest& rq, hipStream_t s,
                        void* scratch, size_t scratch_bytes) {
  const int nv = nv_of(rq);
  PSX_REQUIRE(nv >= 1 && nv <= kMaxV, PSX_ERR_ARGUMENT, "bad number of W spectra");
  PSX_REQUIRE(rq.lmin >= 0 && rq.lmax >= rq.lmin, PSX_ERR_ARGUMENT, "bad lmin/lmax");
  PSX_REQUIRE(rq.LW >= 1 && rq.LW < (1 << 24), PSX_ERR_ARGUMENT, "bad W length");
  PSX_REQUIRE(rq.lmax <= kLmaxCap, PSX_ERR_ARGUMENT, "lmax above 8191 is not supported");
  // l3 never exceeds l1 + l2 <= 2 lmax: longer W spectra are truncated
  ContractRequest rqe = rq;
  rqe.LW = std::min<int64_t>(rq.LW, 2 * (int64_t)rq.lmax + 1);
  return run_contraction_checked(c, rqe, s, scratch, scratch_bytes);
}

// With band_done, the tiles run as nbands launches in l1 bands and after each
// one an event is recorded on s and listed with the band's done_l; the
// covariance flag is then left to the caller.
static int* run_contraction_checked(Ctx* c, const ContractRequest& rq, hipStream_t s,
                                    void* scratch, size_t scratch_bytes, int nbands,
                                    std::vector<std::pair<int, hipEvent_t>>* band_done) {
  const int nv = nv_of(rq);
  const int n = rq.lmax - rq.lmin + 1;
  PSX_REQUIRE(rq.ldo >= n, PSX_ERR_AXES, "leading dimension smaller than lmax-lmin+1");
  const Config cf = pick_config(rq);
  const CachedTiles& tl =
      get_tiles(c, rq.lmin, rq.lmax, (int)rq.LW, cf.R, cf.waves, band_done ? nbands : 1);
  ensure_tables(c, (int)rq.LW + 2 * rq.lmax + 512);
  const Geometry& g = tl.geo;
  PSX_REQUIRE(g.tab_need + kTabOff < c->tables.n, PSX_ERR_ARGUMENT,
              "tile geometry: table overflow");

  // scratch: [record table][err flag]
  const RecLayout rl = rec_layout(rq);
  const size_t recbytes = (size_t)rl.rs * g.wh_len * sizeof(double);
  PSX_REQUIRE(recbytes + 512 <= scratch_bytes, PSX_ERR_BUDGET, "scratch too small");
  double* rec = (double*)scratch;
  int* err = (int*)((char*)scratch + ((recbytes + 255) & ~(size_t)255));

  RecArgs ra{};
  ra.rec = rec;
  for (int k = 0; k < nv; ++k) ra.W[k] = rq.dW[k];
  ra.len = g.wh_len;
  ra.LW = (int)rq.LW;
  ra.nv = rl.nv;
  ra.ne = rl.ne;
  ra.rs = rl.rs;
  ra.ft = rl.ft;
  ra.ft1 = rl.ft1;
  ra.err = rq.epi == EPI_COV ? err : nullptr;
  ra.raw = rq.epi == EPI_RMAX ? 1 : 0;
  {
    dim3 grid((unsigned)((g.wh_len + 255) / 256));
    if (s == c->stream) {
      TimedLaunch tlv(c, "psx_records");
      hipLaunchKernelGGL(record_kernel, grid, dim3(256), 0, s, ra);
    } else {
      hipLaunchKernelGGL(record_kernel, grid, dim3(256), 0, s, ra);
      PSX_HIP(hipGetLastError());
    }
  }

  KArgs ka{};
  ka.tiles = tl.d;
  ka.ntiles = tl.n;
  ka.lmin = rq.lmin;
  ka.lmax = rq.lmax;
  ka.LW = (int)rq.LW;
  ka.G = c->tables.dG;
  ka.Z = c->tables.dZ;
  ka.D = c->tables.dD;
  ka.IG = c->tables.dIG;
  ka.rec = rec;
  ka.ldo = rq.ldo;
  ka.doff = rq.doff;
  if (rq.epi == EPI_MCM || rq.epi == EPI_RMAX) {
    for (int k = 0; k < nv; ++k) ka.out[k] = rq.dout[k];
  } else {
    ka.out[0] = rq.dout[0];
    for (int i = 0; i < 4; ++i) {
      ka.spec[i] = rq.dspec[i];
      ka.ratio[i] = rq.dratio[i];
    }
  }
  ka.err = err;
  const int threads = 64 * cf.waves;
  if (band_done) {
    // every band on s, an event after each (one timed region for all)
    PSX_REQUIRE(s == c->stream && tl.bands.size() <= kMaxBands, PSX_ERR_ARGUMENT,
                "banded contraction needs the context stream");
    ensure_band_events(c);
    TimedLaunch tlch(c, cf.name);
    for (size_t b = 0; b < tl.bands.size(); ++b) {
      const TileBand& tb = tl.bands[b];
      if (tb.nwg > 0) {
        KArgs kb = ka;
        kb.tiles = tl.d + (size_t)tb.wg0 * cf.waves;
        kb.ntiles = tb.nwg * cf.waves;
        hipLaunchKernelGGL(cf.fn, dim3((unsigned)tb.nwg), dim3(threads), 0, s, kb);
        PSX_HIP(hipGetLastError());
      }
      PSX_HIP(hipEventRecord(c->band_ev[b], s));
      band_done->emplace_back(tb.done_l, c->band_ev[b]);
    }
    return rq.epi == EPI_COV ? err : nullptr;
  }
  const unsigned blocks = (unsigned)((tl.n + cf.waves - 1) / cf.waves);
  {
    if (s == c->stream) {
      TimedLaunch tlch(c, cf.name);
      hipLaunchKernelGGL(cf.fn, dim3(blocks), dim3(threads), 0, s, ka);
    } else {
      hipLaunchKernelGGL(cf.fn, dim3(blocks), dim3(threads), 0, s, ka);
      PSX_HIP(hipGetLastError());
    }
  }
  if (rq.epi == EPI_COV && s == c->stream) {
    int h = 0;
    PSX_HIP(hipMemcpyAsync(&h, err, sizeof(int), hipMemcpyDeviceToHost, s));
    PSX_HIP(hipStreamSynchronize(s));
    PSX_REQUIRE(h == 0, PSX_ERR_DOMAIN,
                "sqrt of a negative spectrum product (Julia DomainError)");
  }
  return rq.epi == EPI_COV ? err : nullptr;
}

void run_contraction(Ctx* c, const ContractRequest& rq) {
  void* scratch = ctx_scratch(c, contraction_scratch_bytes(rq));
  run_contraction_on(c, rq, c->stream, scratch, c->scratch_bytes);
}

void run_contraction_to_host(Ctx* c, const ContractRequest& rq, double* const* host,
                             int64_t ldh) {
  const int n = rq.lmax - rq.lmin + 1;
  const int nout = rq.epi == EPI_MCM ? nv_of(rq) : 1;
  PSX_REQUIRE(rq.ldo == n, PSX_ERR_ARGUMENT, "device outputs must be packed (ld = n)");
  for (int k = 0; k < nout; ++k) PSX_REQUIRE(host[k], PSX_ERR_ARGUMENT, "null output pointer");
  // bands only pay when the outputs outgrow one staging buffer
  const bool banded = (size_t)nout * n * (size_t)n * sizeof(double) > kStageBytes;
  void* scratch = ctx_scratch(c, contraction_scratch_bytes(rq));
  std::vector<std::pair<int, hipEvent_t>> done;
  int* err = nullptr;
  if (banded) {
    const int nv = nv_of(rq);
    PSX_REQUIRE(nv >= 1 && nv <= kMaxV, PSX_ERR_ARGUMENT, "bad number of W spectra");
    PSX_REQUIRE(rq.LW >= 1 && rq.LW < (1 << 24), PSX_ERR_ARGUMENT, "bad W length");
    PSX_REQUIRE(rq.lmax <= kLmaxCap, PSX_ERR_ARGUMENT, "lmax above 8191 is not supported");
    ContractRequest rqe = rq;
    rqe.LW = std::min<int64_t>(rq.LW, 2 * (int64_t)rq.lmax + 1);
    err = run_contraction_checked(c, rqe, c->stream, scratch, c->scratch_bytes, kMaxBands,
                                  &done);
  } else {
    // one launch; it checks the covariance flag itself, and the copy follows
    // it on the context stream
    run_contraction_on(c, rq, c->stream, scratch, c->scratch_bytes);
    std::vector<D2HMatrix> whole;
    for (int k = 0; k < nout; ++k) whole.push_back({host[k], ldh, rq.dout[k], n, n, n});
    d2h_matrices(c, whole.data(), nout);
    return;
  }
  // columns [j0, j1) of every output once the band completing them is done
  std::vector<D2HMatrix> pieces;
  int64_t j0 = 0;
  for (const auto& bd : done) {
    const int64_t j1 = std::min<int64_t>(n, bd.first - rq.lmin);
    if (j1 <= j0) continue;
    for (int k = 0; k < nout; ++k)
      pieces.push_back({host[k] + j0 * ldh, ldh, rq.dout[k] + j0 * n, n, n, j1 - j0,
                        bd.second});
    j0 = j1;
  }
  PSX_REQUIRE(j0 == n, PSX_ERR_ARGUMENT, "band bookkeeping left columns uncovered");
  d2h_matrices_overlapped(c, pieces.data(), (int)pieces.size());
  PSX_HIP(hipStreamSynchronize(c->stream));
  if (err) {
    int h = 0;
    PSX_HIP(hipMemcpy(&h, err, sizeof(int), hipMemcpyDeviceToHost));
    PSX_REQUIRE(h == 0, PSX_ERR_DOMAIN,
                "sqrt of a negative spectrum product (Julia DomainError)");
  }
}

// ---- allocation-free mode-coupling path -----------------------------------
// psx_mcm_d_noalloc runs the psx_mcm_d kernels with everything they need
// beyond the caller's W and M in one file-scope device array: the four gamma
// tables (built for lmax <= 8191 and copied in once per device at context
// creation), the launch's tile list (copied in when it changes) and the
// per-call record table.  Host code never passes the array's name to a
// kernel: the device address of each device's copy comes from
// hipGetSymbolAddress, looked up once per device.  Every call on a device
// shares that device's array and overwrites it, so calls that overlap on one
// device (from several contexts or threads) race.
namespace {
constexpr size_t kNoAllocBytes = size_t(4) << 20;
}  // namespace

static __device__ __align__(256) char scratch[kNoAllocBytes];

namespace {

constexpr int kNoAllocDevices = 64;
// largest table index a call with lmax <= kLmaxCap touches (as ensure_tables
// is asked for by run_contraction_checked with LW <= 2 lmax + 1)
constexpr int kNoAllocTabIndex = (2 * kLmaxCap + 1) + 2 * kLmaxCap + 512;

struct NoAllocDevice {
  char* base = nullptr;  // device address of this device's `scratch`
  int ntab = 0;          // entries per gamma table at the front of the array
  bool have_tiles = false;
  TileKey tiles_key{};   // tile list currently in the array
};

NoAllocDevice g_noalloc[kNoAllocDevices];
std::mutex g_noalloc_mu;

struct HostTiles {
  std::vector<Tile> tiles;
  Geometry geo;
};
std::map<TileKey, HostTiles> g_noalloc_tiles;  // host-side tile lists

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

void noalloc_init_device(int device) {
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  PSX_REQUIRE(device >= 0 && device < kNoAllocDevices, PSX_ERR_ARGUMENT,
              "device index beyond the allocation-free path's table");
  NoAllocDevice& d = g_noalloc[device];
  if (d.base) return;
  void* p = nullptr;
  PSX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  PSX_REQUIRE(p != nullptr, PSX_ERR_DEVICE, "hipGetSymbolAddress returned a null address");
  const int n = table_entries(kNoAllocTabIndex);
  const size_t tb = sizeof(double) * (size_t)n;
  PSX_REQUIRE(4 * tb <= kNoAllocBytes, PSX_ERR_DEVICE, "static scratch smaller than its tables");
  std::vector<double> t(4 * (size_t)n);
  build_tables_host(n, t.data(), t.data() + n, t.data() + 2 * (size_t)n,
                    t.data() + 3 * (size_t)n);
  PSX_HIP(hipMemcpy(p, t.data(), 4 * tb, hipMemcpyHostToDevice));
  d.ntab = n;
  d.base = (char*)p;
}

void run_contraction_noalloc(Ctx* c, const ContractRequest& rq0) {
  PSX_REQUIRE(rq0.epi == EPI_MCM && nv_of(rq0) == 1, PSX_ERR_ARGUMENT,
              "the allocation-free path computes one mode-coupling block");
  PSX_REQUIRE(rq0.lmin >= 0 && rq0.lmax >= rq0.lmin, PSX_ERR_ARGUMENT, "bad lmin/lmax");
  PSX_REQUIRE(rq0.lmax <= kLmaxCap, PSX_ERR_ARGUMENT,
              "psx_mcm_d_noalloc: lmax above 8191 is beyond its static scratch");
  PSX_REQUIRE(rq0.LW >= 1 && rq0.LW < (1 << 24), PSX_ERR_ARGUMENT, "bad W length");
  ContractRequest rq = rq0;
  rq.LW = std::min<int64_t>(rq0.LW, 2 * (int64_t)rq0.lmax + 1);
  const int n = rq.lmax - rq.lmin + 1;
  PSX_REQUIRE(rq.ldo >= n, PSX_ERR_AXES, "leading dimension smaller than lmax-lmin+1");
  const Config cf = pick_config(rq);

  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  PSX_REQUIRE(c->device >= 0 && c->device < kNoAllocDevices && g_noalloc[c->device].base,
              PSX_ERR_DEVICE, "static scratch address not set up for this device");
  NoAllocDevice& d = g_noalloc[c->device];
  const TileKey key{rq.lmin, rq.lmax, (int)rq.LW, cf.R, cf.waves, 1};
  auto it = g_noalloc_tiles.find(key);
  if (it == g_noalloc_tiles.end()) {
    HostTiles ht;
    ht.tiles = make_tiles(rq.lmin, rq.lmax, (int)rq.LW, cf.R, cf.waves);
    ht.geo = check_geometry(ht.tiles, rq.lmax, (int)rq.LW, cf.R, cf.waves, 0);
    it = g_noalloc_tiles.emplace(key, std::move(ht)).first;
  }
  const HostTiles& ht = it->second;
  const Geometry& g = ht.geo;
  PSX_REQUIRE(g.tab_need + kTabOff < d.ntab, PSX_ERR_ARGUMENT, "tile geometry: table overflow");

  // [G Z D IG tables][tile list][record table][flag]
  const RecLayout rl = rec_layout(rq);
  const size_t tab_bytes = 4 * sizeof(double) * (size_t)d.ntab;
  const size_t tile_bytes = sizeof(Tile) * ht.tiles.size();
  const size_t rec_bytes = (size_t)rl.rs * g.wh_len * sizeof(double);
  const size_t need = align256(tab_bytes) + align256(tile_bytes) + align256(rec_bytes) + 256;
  PSX_REQUIRE(need <= kNoAllocBytes, PSX_ERR_BUDGET,
              "psx_mcm_d_noalloc: the call needs " + std::to_string(need) +
                  " bytes of the " + std::to_string(kNoAllocBytes) + "-byte static scratch");
  const double* tab = (const double*)d.base;
  Tile* dtiles = (Tile*)(d.base + align256(tab_bytes));
  double* rec = (double*)(d.base + align256(tab_bytes) + align256(tile_bytes));

  if (!d.have_tiles || d.tiles_key < key || key < d.tiles_key) {
    PSX_HIP(hipMemcpyAsync(dtiles, ht.tiles.data(), tile_bytes, hipMemcpyHostToDevice,
                           c->stream));
    d.have_tiles = true;
    d.tiles_key = key;
  }

  RecArgs ra{};
  ra.rec = rec;
  ra.W[0] = rq.dW[0];
  ra.len = g.wh_len;
  ra.LW = (int)rq.LW;
  ra.nv = rl.nv;
  ra.ne = rl.ne;
  ra.rs = rl.rs;
  ra.ft = rl.ft;
  ra.ft1 = rl.ft1;
  ra.err = nullptr;
  hipLaunchKernelGGL(record_kernel, dim3((unsigned)((g.wh_len + 255) / 256)), dim3(256), 0,
                     c->stream, ra);
  PSX_HIP(hipGetLastError());

  KArgs ka{};
  ka.tiles = dtiles;
  ka.ntiles = (int)ht.tiles.size();
  ka.lmin = rq.lmin;
  ka.lmax = rq.lmax;
  ka.LW = (int)rq.LW;
  ka.G = tab;
  ka.Z = tab + d.ntab;
  ka.D = tab + 2 * (size_t)d.ntab;
  ka.IG = tab + 3 * (size_t)d.ntab;
  ka.rec = rec;
  ka.ldo = rq.ldo;
  ka.out[0] = rq.dout[0];
  ka.err = nullptr;
  const unsigned blocks = (unsigned)((ka.ntiles + cf.waves - 1) / cf.waves);
  hipLaunchKernelGGL(cf.fn, dim3(blocks), dim3(64 * cf.waves), 0, c->stream, ka);
  PSX_HIP(hipGetLastError());
}

}  // namespace psx
