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
// odd) every term factorises into a lane factor gamma_{l1-c} gamma_{l2-c}
// (or z_{l1-c-1} z_{l2-c-1}) and a factor that depends on (S, c) only:
//   V_even[c][l3] = What_{l3} gamma_c / ((2(l3+c)+1) gamma_{l3+c})
//   V_odd [c][l3] = What_{l3} z_c / gamma_{l3+c+1},   What = (2l3+1) W / (4 pi)
// V is precomputed once per W spectrum (one HBM pass); the main kernel then
// spends 2 FP64 VALU ops per term (mul + fma) for (0 0 0)^2 and no sqrt, no
// division and no recurrence.  All integer-valued polynomials (P, Q, D0, D1)
// are exact in FP64 for lmax <= ~6000.
//
// Work decomposition: one wavefront owns a tile of 64 consecutive l1 (lanes)
// times R consecutive anti-diagonals S (register slots); for a fixed S the
// lanes walk along an anti-diagonal of the (l1, l2) triangle.  Per step c a
// lane reads gamma_{l1-c} from LDS and slides a register window over
// gamma_{l2-c} (one LDS read per step), while V[c][S-2c] is uniform across
// the wave and is read with scalar loads.
#include "psx_internal.h"

#include <algorithm>
#include <cmath>

namespace psx {

namespace {

constexpr int kVOff = 192;   // V column of l3 = 0
constexpr int kTabLo = -176; // first table index staged into LDS
constexpr int kTabHiPad = 128;
constexpr double kInv4Pi = 0.07957747154594766788444188168625718;  // 1/(4 pi)

struct Tile {
  int x0, S0;
};

constexpr int kMaxV = 12;

struct KArgs {
  const Tile* tiles;
  int ntiles;
  int lmin, lmax;
  int LW;
  const double* G;   // device tables, index m at [m + kTabOff]
  const double* Z;
  const double* T;
  int tab_n;         // entries staged into LDS (indices kTabLo .. kTabLo+tab_n-1)
  const double* V[kMaxV];
  int64_t vpitch;
  double* out[kMaxV];
  int64_t ldo;
  const double* spec[4];
  const double* ratio[4];
  int* err;
};

template <typename T>
__device__ __forceinline__ T sload(const T* p, int64_t i) {
  // Uniform address: force a scalar (SMEM) load through the constant
  // address space so the value lands in SGPRs.
  return ((const __attribute__((address_space(4))) T*)p)[i];
}

__device__ __forceinline__ int ufl(int v) { return __builtin_amdgcn_readfirstlane(v); }

// ---- covariance epilogues (reference src/covariance.jl) -------------------
// xi[] is in family order; see cov_layout() for the W -> xi mapping.
__device__ __forceinline__ double sqc(double x, bool& bad) {
  if (x < 0) bad = true;
  return sqrt(x);
}

template <int COVK>
__device__ __forceinline__ double cov_combine(int a, int b, const double* xi,
                                              const KArgs& A, bool& bad) {
  const double* const* S = A.spec;
  const double* const* R = A.ratio;
  if constexpr (COVK == PSX_COV_TTTT || COVK == PSX_COV_EEEE) {
    // :110-118 / :171-179
    const double *ip = S[0], *jq = S[1], *iq = S[2], *jp = S[3];
    const double *rip = R[0], *rjq = R[1], *riq = R[2], *rjp = R[3];
    return sqc(ip[a] * ip[b] * jq[a] * jq[b], bad) * xi[0] +
           sqc(iq[a] * iq[b] * jp[a] * jp[b], bad) * xi[1] +
           sqc(ip[a] * ip[b], bad) * xi[2] * rjq[a] * rjq[b] +
           sqc(jq[a] * jq[b], bad) * xi[3] * rip[a] * rip[b] +
           sqc(iq[a] * iq[b], bad) * xi[4] * rjp[a] * rjp[b] +
           sqc(jp[a] * jp[b], bad) * xi[5] * riq[a] * riq[b] +
           xi[6] * rip[a] * rjq[a] * rip[b] * rjq[b] +
           xi[7] * riq[a] * rjp[a] * riq[b] * rjp[b];
  } else if constexpr (COVK == PSX_COV_TTTE) {
    // :225-230
    const double *TTip = S[0], *TTjp = S[1], *TEiq = S[2], *TEjq = S[3];
    const double *rip = R[0], *rjp = R[1];
    return (sqc(TTip[a] * TTip[b], bad) * (TEjq[a] + TEjq[b]) * xi[0] +
            sqc(TTjp[a] * TTjp[b], bad) * (TEiq[a] + TEiq[b]) * xi[1] +
            (TEjq[a] + TEjq[b]) * xi[2] * rip[a] * rip[b] +
            (TEiq[a] + TEiq[b]) * xi[3] * rjp[a] * rjp[b]) /
           2;
  } else if constexpr (COVK == PSX_COV_TETE) {
    // :292-297; xi[0] = Ξ_TT(W2), xi[1..4] = Ξ_TE(W1, W3, W4, W5)
    const double *TTip = S[0], *EEjq = S[1], *TEiq = S[2], *TEjp = S[3];
    const double *rTTip = R[0], *rPPjq = R[1];
    return sqc(TTip[a] * TTip[b] * EEjq[a] * EEjq[b], bad) * xi[1] +
           0.5 * (TEiq[a] * TEjp[b] + TEjp[a] * TEiq[b]) * xi[0] +
           sqc(TTip[a] * TTip[b], bad) * xi[2] * rPPjq[a] * rPPjq[b] +
           sqc(EEjq[a] * EEjq[b], bad) * xi[3] * rTTip[a] * rTTip[b] +
           xi[4] * rTTip[a] * rTTip[b] * rPPjq[a] * rPPjq[b];
  } else if constexpr (COVK == PSX_COV_TEEE || COVK == PSX_COV_TEEE_PLANCK) {
    // :362-367 / :392-397
    const double *EEjq = S[0], *EEjp = S[1], *TEip = S[2], *TEiq = S[3];
    const double *rjq = R[0], *rjp = R[1];
    return (sqc(EEjq[a] * EEjq[b], bad) * (TEip[a] + TEip[b]) * xi[0] +
            sqc(EEjp[a] * EEjp[b], bad) * (TEiq[a] + TEiq[b]) * xi[1] +
            (TEip[a] + TEip[b]) * xi[2] * rjq[a] * rjq[b] +
            (TEiq[a] + TEiq[b]) * xi[3] * rjp[a] * rjp[b]) /
           2;
  } else {  // PSX_COV_TTEE, :438-441
    const double *TEip = S[0], *TEiq = S[1], *TEjq = S[2], *TEjp = S[3];
    return ((TEip[a] * TEjq[b] + TEjq[a] * TEip[b]) * xi[0] +
            (TEiq[a] * TEjp[b] + TEjp[a] * TEiq[b]) * xi[1]) /
           2;
  }
}

// ---- the contraction kernel ---------------------------------------------
// EPI: EPI_MCM or EPI_COV; COVK: covariance kind for EPI_COV.
template <int R, int N00, int N02, int NPE, int NPO, int EPI, int COVK>
__global__ void __launch_bounds__(256) contract_kernel(KArgs A) {
  constexpr int NV = N00 + N02 + NPE + NPO;
  constexpr bool kG = (N00 + N02 + NPE) > 0;
  constexpr bool kZ = NPO > 0;
  constexpr bool kE = (N02 + NPE) > 0;
  extern __shared__ double lds[];
  double* sG = lds;
  double* sZ = lds + (kG ? A.tab_n : 0);

  // Stage the gamma / z tables for indices kTabLo .. kTabLo + tab_n - 1.
  for (int i = threadIdx.x; i < A.tab_n; i += blockDim.x) {
    if (kG) sG[i] = A.G[i + kTabLo + kTabOff];
    if (kZ) sZ[i] = A.Z[i + kTabLo + kTabOff];
  }
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int wave = ufl(threadIdx.x >> 6);
  const int tile = ufl(blockIdx.x * (blockDim.x >> 6) + wave);
  if (tile >= A.ntiles) return;
  const int x0 = sload(&A.tiles[tile].x0, 0);
  const int S0 = sload(&A.tiles[tile].S0, 0);
  const int lmax = A.lmax;
  const int l1 = x0 + lane;

  // summation range over c, rounded out to whole blocks of R
  const int chi = min(x0 + 63, lmax);
  const int clo = max(0, (S0 - A.LW + 1) >> 1);
  const int cbeg = ufl((clo / R) * R);
  const int nblk = ufl(max(0, (chi + 1 - cbeg + R - 1) / R));

  // lane constants per slot
  double sig1[kE ? R : 1], kap[kE ? R : 1], sig2[kZ ? R : 1];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const double L1 = (double)l1, L2 = (double)(S0 + r - l1);
    const double P1 = L1 * (L1 + 1.0), P2 = L2 * (L2 + 1.0);
    if constexpr (kE) {
      sig1[r] = P1 + P2 - 1.0;
      kap[r] = 1.0 + 2.0 * P1 * P2;
    }
    if constexpr (kZ) sig2[r] = P1 + P2 - 2.0;
  }

  double acc[NV][R];
#pragma unroll
  for (int k = 0; k < NV; ++k)
#pragma unroll
    for (int r = 0; r < R; ++r) acc[k][r] = 0.0;

  // sliding windows: slot r at step c needs G[t + r - c], Z[t + r - c - 1]
  // with t = S0 - x0 - lane; held at gw[(r - c) mod R] (cbeg % R == 0).
  const int t = S0 - x0 - lane;
  double gw[kG ? R : 1], zw[kZ ? R : 1];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if constexpr (kG) gw[r] = sG[t + r - cbeg - kTabLo];
    if constexpr (kZ) zw[r] = sZ[t + r - cbeg - 1 - kTabLo];
  }

  const int64_t P = A.vpitch;
  for (int blk = 0; blk < nblk; ++blk) {
    const int c0 = cbeg + blk * R;
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const int c = c0 + u;
      double g1 = 0.0, z1 = 0.0;
      if constexpr (kG) g1 = sG[l1 - c - kTabLo];
      if constexpr (kZ) z1 = sZ[l1 - c - 1 - kTabLo];
      // uniform offsets of V row c at column l3 = S0 - 2c (+r)
      const int64_t voff = (int64_t)c * P + kVOff + S0 - 2 * c;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int wi = ((r - u) % R + R) % R;
        double p = 0.0, Q = 0.0;
        if constexpr (kG) p = g1 * gw[wi];
        if constexpr (kE) {
          const double Te = sload(A.T, (int64_t)kTabOff + S0 + r - 2 * c);
          const double E = sig1[r] - Te;
          Q = fma(E, E, -kap[r]);
        }
#pragma unroll
        for (int k = 0; k < N00; ++k)
          acc[k][r] = fma(sload(A.V[k], voff + r), p, acc[k][r]);
        double pq = 0.0;
        if constexpr (N02 > 0) {
          pq = p * Q;
#pragma unroll
          for (int k = 0; k < N02; ++k)
            acc[N00 + k][r] = fma(sload(A.V[N00 + k], voff + r), pq, acc[N00 + k][r]);
        }
        if constexpr (NPE > 0) {
          const double pqq = (N02 > 0 ? pq : p * Q) * Q;
#pragma unroll
          for (int k = 0; k < NPE; ++k)
            acc[N00 + N02 + k][r] =
                fma(sload(A.V[N00 + N02 + k], voff + r), pqq, acc[N00 + N02 + k][r]);
        }
        if constexpr (NPO > 0) {
          const double To = sload(A.T, (int64_t)kTabOff + S0 + r - 2 * c - 1);
          const double D1 = sig2[r] - To;
          const double pd = (z1 * zw[wi]) * (D1 * D1);
#pragma unroll
          for (int k = 0; k < NPO; ++k)
            acc[N00 + N02 + NPE + k][r] = fma(sload(A.V[N00 + N02 + NPE + k], voff - 1 + r),
                                              pd, acc[N00 + N02 + NPE + k][r]);
        }
      }
      // advance the windows to step c + 1: the slot that held r = R-1
      // receives G[t - c - 1]
      const int ni = (R - 1 - u) % R;
      if constexpr (kG) gw[ni] = sG[t - c - 1 - kTabLo];
      if constexpr (kZ) zw[ni] = sZ[t - c - 2 - kTabLo];
    }
  }

  // ---- epilogue ----
  const int lmin = A.lmin;
  const double L1 = (double)l1;
  const double pi1 = (L1 - 1.0) * L1 * (L1 + 1.0) * (L1 + 2.0);
  bool bad = false;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int l2 = S0 + r - l1;
    if (l1 < lmin || l1 > l2 || l2 > lmax) continue;
    const double L2 = (double)l2;
    const double pi2 = (L2 - 1.0) * L2 * (L2 + 1.0) * (L2 + 2.0);
    const bool spin2 = (l1 >= 2) && (l2 >= 2);
    const double k22 = spin2 ? 1.0 / (4.0 * pi1 * pi2) : 0.0;
    const double k02 = spin2 ? 1.0 / (2.0 * sqrt(pi1 * pi2)) : 0.0;
    double xi[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const double kap_k = k < N00 ? 1.0 : (k < N00 + N02 ? k02 : k22);
      xi[k] = kap_k * acc[k][r];
    }
    const int64_t i12 = (int64_t)(l1 - lmin) + (int64_t)(l2 - lmin) * A.ldo;
    const int64_t i21 = (int64_t)(l2 - lmin) + (int64_t)(l1 - lmin) * A.ldo;
    if constexpr (EPI == EPI_MCM) {
#pragma unroll
      for (int k = 0; k < NV; ++k) {
        A.out[k][i12] = (2.0 * L2 + 1.0) * xi[k];
        A.out[k][i21] = (2.0 * L1 + 1.0) * xi[k];
      }
    } else {
      const double v = cov_combine<COVK>(l1, l2, xi, A, bad);
      A.out[0][i12] = v;
      A.out[0][i21] = v;
    }
  }
  if constexpr (EPI == EPI_COV) {
    if (bad) atomicOr(A.err, 1);
  }
}

// ---- V table construction ------------------------------------------------
struct VArgs {
  double* V[kMaxV];
  const double* W[kMaxV];
  int odd[kMaxV];
  int64_t pitch;
  int LW;
  const double* G;
  const double* Z;
  const double* D;
  const double* IG;
};

__global__ void __launch_bounds__(256) vtable_kernel(VArgs a) {
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  const int k = blockIdx.z;
  if (col >= a.pitch) return;
  const int l3 = (int)col - kVOff;
  double v = 0.0;
  if (l3 >= 0 && l3 < a.LW) {
    const double wh = (2.0 * l3 + 1.0) * a.W[k][l3] * kInv4Pi;
    if (a.odd[k])
      v = wh * a.Z[c + kTabOff] * a.IG[l3 + c + 1 + kTabOff];
    else
      v = wh * a.G[c + kTabOff] * a.D[l3 + c + kTabOff];
  }
  a.V[k][(int64_t)c * a.pitch + col] = v;
}

// ---- tile lists ----------------------------------------------------------
struct TileKey {
  int lmin, lmax, LW, R;
  bool operator<(const TileKey& o) const {
    return std::tie(lmin, lmax, LW, R) < std::tie(o.lmin, o.lmax, o.LW, o.R);
  }
};

std::vector<Tile> make_tiles(int lmin, int lmax, int LW, int R) {
  struct TC {
    Tile t;
    int cost;
  };
  std::vector<TC> v;
  for (int x0 = lmin; x0 <= lmax; x0 += 64) {
    const int l1hi = std::min(x0 + 63, lmax);
    const int Slo = 2 * x0, Shi = l1hi + lmax;
    for (int S0 = Slo; S0 <= Shi; S0 += R) {
      const int clo = std::max(0, (S0 - LW + 1) >> 1);
      const int cbeg = (clo / R) * R;
      const int nblk = std::max(0, (l1hi + 1 - cbeg + R - 1) / R);
      v.push_back({{x0, S0}, nblk});
    }
  }
  std::stable_sort(v.begin(), v.end(),
                   [](const TC& a, const TC& b) { return a.cost > b.cost; });
  std::vector<Tile> out;
  out.reserve(v.size());
  for (auto& e : v) out.push_back(e.t);
  return out;
}

struct Geometry {
  int R;
  int64_t pitch;  // V row pitch (doubles)
  int nrows;      // V rows
  int tab_n;      // LDS table entries
  int tab_need;   // largest table index any read touches
};

// Index ranges touched by contract_kernel for this tile set, checked against
// the buffers before launch (see the access list in contract_kernel).
Geometry check_geometry(const std::vector<Tile>& tiles, int lmax, int LW, int R,
                        int tables_n) {
  Geometry g{};
  g.R = R;
  int tmin = 0, tmax = 0, vmin = 0, vmax = 0, rmax = 0, Tmin = 0, Tmax = 0;
  bool first = true;
  for (const Tile& t : tiles) {
    const int chi = std::min(t.x0 + 63, lmax);
    const int clo = std::max(0, (t.S0 - LW + 1) >> 1);
    const int cbeg = (clo / R) * R;
    const int nblk = std::max(0, (chi + 1 - cbeg + R - 1) / R);
    const int cend = cbeg + nblk * R;  // exclusive
    // LDS table reads (G and Z, incl. the -1 / -2 shifts; window start and
    // the loads for steps cbeg .. cend)
    const int lo = std::min({t.x0 - cend, t.S0 - t.x0 - 63 - cend - 1,
                             t.S0 - t.x0 - 63 - cbeg - 1});
    const int hi = std::max({t.x0 + 63 - cbeg, t.S0 - t.x0 + R - 1 - cbeg});
    // V columns (l3) and rows, T table
    const int l3lo = t.S0 - 2 * (cend - 1) - 1;
    const int l3hi = t.S0 + R - 1 - 2 * cbeg;
    if (first) {
      tmin = lo; tmax = hi; vmin = l3lo; vmax = l3hi; rmax = cend - 1;
      first = false;
    } else {
      tmin = std::min(tmin, lo); tmax = std::max(tmax, hi);
      vmin = std::min(vmin, l3lo); vmax = std::max(vmax, l3hi);
      rmax = std::max(rmax, cend - 1);
    }
  }
  Tmin = vmin;
  Tmax = vmax;
  PSX_REQUIRE(tmin >= kTabLo, PSX_ERR_ARGUMENT, "tile geometry: LDS table underflow");
  g.tab_n = tmax - kTabLo + 1;
  PSX_REQUIRE(vmin >= -kVOff, PSX_ERR_ARGUMENT, "tile geometry: V column underflow");
  g.pitch = ((int64_t)kVOff + std::max(vmax, LW) + 1 + 15) & ~(int64_t)15;
  g.nrows = rmax + 1;
  PSX_REQUIRE(Tmin >= -kTabOff, PSX_ERR_ARGUMENT, "tile geometry: T table underflow");
  // V construction reads D/IG up to (l3 + c + 1) with l3 < LW, c < nrows
  g.tab_need = std::max({Tmax, g.tab_n + kTabLo, LW + g.nrows + 2, tmax});
  PSX_REQUIRE(g.tab_need + kTabOff < tables_n || tables_n == 0, PSX_ERR_ARGUMENT,
              "tile geometry: table overflow");
  return g;
}

struct CachedTiles {
  Tile* d = nullptr;
  int n = 0;
  std::vector<Tile> host;
};

std::map<std::pair<void*, TileKey>, CachedTiles> g_tile_cache;
std::mutex g_tile_mu;

const CachedTiles& get_tiles(Ctx* c, int lmin, int lmax, int LW, int R) {
  std::lock_guard<std::mutex> lk(g_tile_mu);
  auto key = std::make_pair((void*)c, TileKey{lmin, lmax, LW, R});
  auto it = g_tile_cache.find(key);
  if (it != g_tile_cache.end()) return it->second;
  CachedTiles ct;
  ct.host = make_tiles(lmin, lmax, LW, R);
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

template <int R, int N00, int N02, int NPE, int NPO, int EPI, int COVK>
Config cfg(int waves, const char* name) {
  return Config{R, waves, contract_kernel<R, N00, N02, NPE, NPO, EPI, COVK>, name};
}

Config pick_config(const ContractRequest& rq) {
  const int* f = rq.nfam;
  if (rq.epi == EPI_MCM) {
    if (f[0] == 1 && f[1] == 0 && f[2] == 0 && f[3] == 0)
      return cfg<16, 1, 0, 0, 0, EPI_MCM, 0>(4, "psx_mcm_00");
    if (f[0] == 0 && f[1] == 1 && f[2] == 0 && f[3] == 0)
      return cfg<8, 0, 1, 0, 0, EPI_MCM, 0>(4, "psx_mcm_02");
    if (f[0] == 0 && f[1] == 0 && f[2] == 1 && f[3] == 0)
      return cfg<8, 0, 0, 1, 0, EPI_MCM, 0>(4, "psx_mcm_pp");
    if (f[0] == 0 && f[1] == 0 && f[2] == 0 && f[3] == 1)
      return cfg<8, 0, 0, 0, 1, EPI_MCM, 0>(4, "psx_mcm_mm");
    if (f[0] == 0 && f[1] == 0 && f[2] == 1 && f[3] == 1)
      return cfg<8, 0, 0, 1, 1, EPI_MCM, 0>(4, "psx_mcm_pm");
    if (f[0] == 1 && f[1] == 2 && f[2] == 1 && f[3] == 1)
      return cfg<4, 1, 2, 1, 1, EPI_MCM, 0>(4, "psx_mcm_set");
    if (f[0] == 1 && f[1] == 1 && f[2] == 1 && f[3] == 1)
      return cfg<4, 1, 1, 1, 1, EPI_MCM, 0>(4, "psx_mcm_set1");
  } else {
    switch (rq.cov_kind) {
      case PSX_COV_TTTT: return cfg<4, 8, 0, 0, 0, EPI_COV, PSX_COV_TTTT>(4, "psx_cov_tttt");
      case PSX_COV_EEEE: return cfg<4, 0, 0, 8, 0, EPI_COV, PSX_COV_EEEE>(4, "psx_cov_eeee");
      case PSX_COV_TTTE: return cfg<8, 4, 0, 0, 0, EPI_COV, PSX_COV_TTTE>(4, "psx_cov_ttte");
      case PSX_COV_TETE: return cfg<4, 1, 4, 0, 0, EPI_COV, PSX_COV_TETE>(4, "psx_cov_tete");
      case PSX_COV_TEEE: return cfg<4, 0, 4, 0, 0, EPI_COV, PSX_COV_TEEE>(4, "psx_cov_teee");
      case PSX_COV_TEEE_PLANCK:
        return cfg<4, 0, 0, 4, 0, EPI_COV, PSX_COV_TEEE_PLANCK>(4, "psx_cov_teee_planck");
      case PSX_COV_TTEE: return cfg<8, 2, 0, 0, 0, EPI_COV, PSX_COV_TTEE>(4, "psx_cov_ttee");
    }
  }
  throw Error{PSX_ERR_ARGUMENT, "unsupported contraction configuration"};
}

int nv_of(const ContractRequest& rq) {
  return rq.nfam[0] + rq.nfam[1] + rq.nfam[2] + rq.nfam[3];
}

}  // namespace

static void run_contraction_checked(Ctx* c, const ContractRequest& rq, hipStream_t s,
                                    void* scratch, size_t scratch_bytes);

size_t contraction_scratch_bytes(const ContractRequest& rq) {
  // V tables dominate: nv * nrows * pitch doubles (+ error flag)
  const int R = 16;
  const int64_t nrows = rq.lmax + 64 + R + 1;
  const int64_t LW = std::min<int64_t>(rq.LW, 2 * (int64_t)rq.lmax + 1);
  const int64_t pitch = kVOff + std::max<int64_t>(LW + 3 * R, rq.lmax + 64) + 32;
  return (size_t)nv_of(rq) * nrows * pitch * sizeof(double) + 256;
}

void run_contraction_on(Ctx* c, const ContractRequest& rq, hipStream_t s,
                        void* scratch, size_t scratch_bytes) {
  const int nv = nv_of(rq);
  PSX_REQUIRE(nv >= 1 && nv <= kMaxV, PSX_ERR_ARGUMENT, "bad number of W spectra");
  PSX_REQUIRE(rq.lmin >= 0 && rq.lmax >= rq.lmin, PSX_ERR_ARGUMENT, "bad lmin/lmax");
  PSX_REQUIRE(rq.LW >= 1 && rq.LW < (1 << 24), PSX_ERR_ARGUMENT, "bad W length");
  PSX_REQUIRE(rq.lmax <= 16383, PSX_ERR_ARGUMENT, "lmax above 16383 is not supported");
  // l3 never exceeds l1 + l2 <= 2 lmax: longer W spectra are truncated
  ContractRequest rqe = rq;
  rqe.LW = std::min<int64_t>(rq.LW, 2 * (int64_t)rq.lmax + 1);
  run_contraction_checked(c, rqe, s, scratch, scratch_bytes);
}

static void run_contraction_checked(Ctx* c, const ContractRequest& rq, hipStream_t s,
                                    void* scratch, size_t scratch_bytes) {
  const int nv = nv_of(rq);
  const int n = rq.lmax - rq.lmin + 1;
  PSX_REQUIRE(rq.ldo >= n, PSX_ERR_AXES, "leading dimension smaller than lmax-lmin+1");
  const Config cf = pick_config(rq);
  const CachedTiles& tl = get_tiles(c, rq.lmin, rq.lmax, (int)rq.LW, cf.R);
  ensure_tables(c, (int)rq.LW + 2 * rq.lmax + 512);
  const Geometry g = check_geometry(tl.host, rq.lmax, (int)rq.LW, cf.R, c->tables.n);

  // scratch: [V tables][err flag]
  const size_t vbytes = (size_t)nv * g.nrows * g.pitch * sizeof(double);
  PSX_REQUIRE(vbytes + 256 <= scratch_bytes, PSX_ERR_BUDGET, "scratch too small");
  double* V = (double*)scratch;
  int* err = (int*)((char*)scratch + ((vbytes + 255) & ~(size_t)255));

  VArgs va{};
  int fam_of[kMaxV];
  {
    int k = 0;
    for (int fam = 0; fam < 4; ++fam)
      for (int i = 0; i < rq.nfam[fam]; ++i) fam_of[k++] = fam;
  }
  for (int k = 0; k < nv; ++k) {
    va.V[k] = V + (size_t)k * g.nrows * g.pitch;
    va.W[k] = rq.dW[k];
    va.odd[k] = fam_of[k] == FAM_PO;
  }
  va.pitch = g.pitch;
  va.LW = (int)rq.LW;
  va.G = c->tables.dG;
  va.Z = c->tables.dZ;
  va.D = c->tables.dD;
  va.IG = c->tables.dIG;
  {
    dim3 grid((unsigned)((g.pitch + 255) / 256), (unsigned)g.nrows, (unsigned)nv);
    hipLaunchKernelGGL(vtable_kernel, grid, dim3(256), 0, s, va);
    PSX_HIP(hipGetLastError());
  }
  if (rq.epi == EPI_COV) PSX_HIP(hipMemsetAsync(err, 0, sizeof(int), s));

  KArgs ka{};
  ka.tiles = tl.d;
  ka.ntiles = tl.n;
  ka.lmin = rq.lmin;
  ka.lmax = rq.lmax;
  ka.LW = (int)rq.LW;
  ka.G = c->tables.dG;
  ka.Z = c->tables.dZ;
  ka.T = c->tables.dT;
  ka.tab_n = g.tab_n;
  for (int k = 0; k < nv; ++k) ka.V[k] = va.V[k];
  ka.vpitch = g.pitch;
  ka.ldo = rq.ldo;
  if (rq.epi == EPI_MCM) {
    for (int k = 0; k < nv; ++k) ka.out[k] = rq.dout[k];
  } else {
    ka.out[0] = rq.dout[0];
    for (int i = 0; i < 4; ++i) {
      ka.spec[i] = rq.dspec[i];
      ka.ratio[i] = rq.dratio[i];
    }
  }
  ka.err = err;
  const bool kG = (rq.nfam[0] + rq.nfam[1] + rq.nfam[2]) > 0;
  const bool kZ = rq.nfam[3] > 0;
  const size_t lds = (size_t)((kG ? 1 : 0) + (kZ ? 1 : 0)) * g.tab_n * sizeof(double);
  PSX_REQUIRE(lds <= 160 * 1024, PSX_ERR_ARGUMENT, "lmax too large for LDS staging");
  PSX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(cf.fn),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const int threads = 64 * cf.waves;
  const unsigned blocks = (unsigned)((tl.n + cf.waves - 1) / cf.waves);
  {
    if (s == c->stream) {
      TimedLaunch tlch(c, cf.name);
      hipLaunchKernelGGL(cf.fn, dim3(blocks), dim3(threads), lds, s, ka);
    } else {
      hipLaunchKernelGGL(cf.fn, dim3(blocks), dim3(threads), lds, s, ka);
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
}

void run_contraction(Ctx* c, const ContractRequest& rq) {
  void* scratch = ctx_scratch(c, contraction_scratch_bytes(rq));
  run_contraction_on(c, rq, c->stream, scratch, c->scratch_bytes);
}

}  // namespace psx
