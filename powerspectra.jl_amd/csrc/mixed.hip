// Many-W contraction on the matrix cores: K mask spectra contracted against
// one family of 3j products in one pass, for batches of mode-coupling blocks
// (K blocks of one kind from K W spectra) and of covariance blocks (nblk
// blocks of one kind, nW W spectra each).  The reference contracts each W
// separately inside its per-pair loops (src/modecoupling.jl:3-66,78-159;
// src/covariance.jl:92-446); here every family product is computed once and
// multiplied into all K spectra as a matrix product.
//
// Per anti-diagonal S = l1 + l2 the contraction is a true GEMM:
//   X_k(l1, S) = sum_c A(l1, c) B_k(l3),   l3 = S - 2c (- 1 for the odd family)
// with A the family product of the closed forms in contract.hip and
// B_k = What_k = (2 l3 + 1) W_k(l3) / (4 pi).  Rows are 16 consecutive l1
// (one MFMA row group), the inner index is c, columns are the K spectra.
//
// Precision PSX_MIXED runs the product on v_mfma_i32_16x16x64_i8 with an
// error-free split of both operands into balanced base-256 digits
// (d_j in [-128, 127], x = sum_j d_j 256^j exactly):
//   * a power-of-two envelope env(l3) >= max_k |What_k(l3)| / c_k (c_k a
//     power of two per spectrum) is moved from B into A:
//     A' = A env,  B'_k = What_k / (c_k env) in [-1, 1],
//     so B' has no dynamic range left across l3 and takes one fixed-point
//     scale 2^54 for all l3 (7 digits: a spectrum far below the envelope
//     at some l3 keeps its relative precision there);
//   * A' takes one exponent per row over its whole family: the exact maximum
//     of |A'| over the row, scaled to < 2^46 (6 digits, split on the fly).
//     The maxima come from the FP64 contraction kernel run as a row-maximum
//     pass (contract.hip, EPI_RMAX: the same family product times env, few
//     ulp apart from this kernel's; the split tolerates |x| < 2^47);
//   * the 27 digit products with s + t >= 5 accumulate in int32 (exact:
//     below 2^30 per diagonal at lmax 8191) over the whole family, seven
//     diagonals per output;
//   * the diagonals are combined in FP64 at the end.
// The dropped products (s + t <= 4) and the rounding of A' to 46 and B' to
// 54 bits put the error near 1e-14 of the element's terms (emulated against
// long double on mask-like spectra before this kernel was written: 15
// products of five digits each left 6e-10 of the element, 21 of six 3e-12;
// with six digits for both, a TE block whose terms cancel 1400x still lost
// 1e-12 of its terms to B' quantised against the envelope; measured per
// block against the FP64 kernels and the oracle by the tests).
//
// Schedule: one wave per 16-row tile of one anti-diagonal S; a workgroup holds
// the same row chunk of S, S + 2, S + 4, S + 6.  Their B rows coincide (B
// depends on l3 = 2h + p only), so the four waves load one B stream
// cooperatively, straight into LDS (global_load_lds), one K-step (64 h)
// ahead into the other of two buffers, with one workgroup barrier per step;
// the lane-factor windows of the family product are loaded one step ahead
// into registers and not touched before their step.  Two workgroups per CU
// (register budget of two waves per SIMD) let one wave's VALU work (the family
// product, its digit split) run beside the other's MFMAs.
//
// Precision PSX_FP64_MFMA runs the same tiles on v_mfma_f64_16x16x4_f64 with
// A and B in FP64 (no splitting): exact FP64 products, a different
// summation order than the VALU kernels.
#include "cov_epi.h"
#include "psx_internal.h"

#include <algorithm>
#include <cmath>

namespace psx {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kMaxCols = 64;     // W spectra in one pass
constexpr int kMaxBlk = 16;      // covariance blocks in one pass
constexpr int kNd = 6;           // int8 digits of A' (split on the fly)
constexpr int kNdB = 7;          // int8 digits of B' (precomputed per call)
constexpr int kDmin = 5;         // digit products with s + t >= kDmin are kept
constexpr int kNdiag = (kNd - 1) + (kNdB - 1) - kDmin + 1;  // diagonals s + t = 5 .. 11
constexpr int kFix = 46;         // fixed-point bits of A' (|x| <= 2^46)
constexpr int kFixB = 54;        // fixed-point bits of B' (|x| <= 2^54)
// 1.5 * 2^52 rounds a double to an integer in the low mantissa bits;
// 0x808080808080 biases the 48-bit integer so that its bytes are the
// balanced digits plus 128
constexpr double kMagic = 6755399441055744.0 + (double)0x808080808080ull;
constexpr double kInv4PiM = 0.07957747154594766788444188168625718;

enum MixMode : int { MIX_I8 = 0, MIX_F64 = 1 };

struct MTile {
  int S, x0, lo, hi;  // anti-diagonal, first row of the row groups, owned rows
  int hb0, nsteps;    // the workgroup's common K-steps: h = hb0 .. hb0 + 64 nsteps - 1
};

struct MixArgs {
  const MTile* tiles;
  int ntiles;
  int lmin, lmax, LW;
  const double* G;
  const double* Z;
  const double* D;
  const double* IG;
  const double* env;        // env[l3] for l3 < envlen
  const v4i* bdig;          // MIX_I8: digits, [t][p][h/16][kpad] x 16 bytes
  const double* bf64;       // MIX_F64: What' table, [p][h][kpad]
  int hblocks;              // 16-h blocks per (digit, parity)
  int kpad;                 // columns padded to a multiple of 16
  int ncol;                 // K
  const double* colscale;   // c_k
  double* const* out;       // device table: EPI_MCM K outputs, EPI_COV nblk outputs
  int64_t ldo;
  int nW;                   // covariance: W spectra per block
  const double* const* spec;   // device table [blk][4]
  const double* const* ratio;  // device table [blk][4]
  int* err;
  // EPI_MCM: the values kappa Xi go to an anti-diagonal-major stage,
  // dmat[(doff[S - 2 lmin] + l1 - lo(S)) kpad + k], lo(S) = max(lmin, S - lmax),
  // and mix_scatter_kernel writes the K outputs from it
  double* dmat;
  const int64_t* doff;
  // MIX_I8: max over its family of |A'| per pair (l1 <= l2), anti-diagonal-major
  // like the stage (the row-maximum pass of the FP64 contraction kernel)
  const double* rowmax;
};

__device__ __forceinline__ int ufl(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Family product (without the kappa factor of the epilogue) at row l1,
// anti-diagonal S and summation index c from the staged factors:
// u1 = T[l1 - c (- 1)], u2 = T[l2 - c (- 1)], v = (row-independent part) env.
// The spin-2 polynomials (exact integers for lmax <= 8191):
//   E = P1 + P2 - 1 - l3(l3 + 1),  Q = E^2 - kappa  (02: p Q, ++: p Q^2)
//   D1 = P1 + P2 - l3 (l3 + 1) - 2  (--: z-product D1^2, l3 odd)
template <int FAM>
__device__ __forceinline__ double fam_value(double u1, double u2, double v, double sig,
                                            double kap, int l3) {
  const double p = (u1 * u2) * v;
  if constexpr (FAM == FAM_00) {
    return p;
  } else if constexpr (FAM == FAM_02 || FAM == FAM_PE) {
    const double T = (double)l3 * (double)(l3 + 1);
    const double E = sig - T;
    const double Q = fma(E, E, -kap);
    return FAM == FAM_02 ? p * Q : (p * Q) * Q;
  } else {
    // l3 is the odd index here: D1 = P1 + P2 - l3 (l3 + 1) - 2
    const double T1 = (double)l3 * (double)(l3 + 1) + 1.0;
    const double D1 = sig - T1;
    return p * (D1 * D1);
  }
}

// bytes of the biased 48-bit integer of x (|x| <= 2^46)
__device__ __forceinline__ void split_words(double x, uint32_t& lo, uint32_t& hi) {
  const double y = x + kMagic;
  const unsigned long long b = (unsigned long long)__double_as_longlong(y);
  lo = (uint32_t)b;
  hi = (uint32_t)(b >> 32);
}

// Digit words of four values a, b, c, d (their split words): word t holds
// digit t of a, b, c, d in bytes 0..3, as signed bytes.
__device__ __forceinline__ void transpose4(uint32_t la, uint32_t lb, uint32_t lc, uint32_t ld,
                                           uint32_t ha, uint32_t hb, uint32_t hc, uint32_t hd,
                                           uint32_t w[kNd]) {
  // v_perm_b32(s0, s1, sel): byte k of the result = byte sel_k of {s0:s1}
  // (bytes 0-3 from s1, 4-7 from s0)
  const uint32_t ab_lo = __builtin_amdgcn_perm(lb, la, 0x05010400u);  // a0 b0 a1 b1
  const uint32_t ab_hi = __builtin_amdgcn_perm(lb, la, 0x07030602u);  // a2 b2 a3 b3
  const uint32_t cd_lo = __builtin_amdgcn_perm(ld, lc, 0x05010400u);
  const uint32_t cd_hi = __builtin_amdgcn_perm(ld, lc, 0x07030602u);
  w[0] = __builtin_amdgcn_perm(cd_lo, ab_lo, 0x05040100u) ^ 0x80808080u;  // a0 b0 c0 d0
  w[1] = __builtin_amdgcn_perm(cd_lo, ab_lo, 0x07060302u) ^ 0x80808080u;  // a1 b1 c1 d1
  w[2] = __builtin_amdgcn_perm(cd_hi, ab_hi, 0x05040100u) ^ 0x80808080u;
  w[3] = __builtin_amdgcn_perm(cd_hi, ab_hi, 0x07060302u) ^ 0x80808080u;
  const uint32_t h_ab = __builtin_amdgcn_perm(hb, ha, 0x05010400u);  // a4 b4 a5 b5
  const uint32_t h_cd = __builtin_amdgcn_perm(hd, hc, 0x05010400u);
  w[4] = __builtin_amdgcn_perm(h_cd, h_ab, 0x05040100u) ^ 0x80808080u;
  w[5] = __builtin_amdgcn_perm(h_cd, h_ab, 0x07060302u) ^ 0x80808080u;
}

// ---- B side: column scales, envelope, digits ------------------------------
struct PrepArgs {
  const double* const* W;   // device table of K column sources
  int K, kpad, LW;
  double* colscale;
  double* env;
  int envlen;
  int8_t* bdig;             // MIX_I8
  double* bf64;             // MIX_F64
  int hblocks;
  int* err;                 // covariance flag, cleared here
};

__device__ __forceinline__ double what_m(const double* W, int l3, int LW) {
  return (l3 >= 0 && l3 < LW) ? (2.0 * l3 + 1.0) * W[l3] * kInv4PiM : 0.0;
}

// c_k = smallest power of two >= max_l3 |What_k(l3)| (1 for an all-zero column)
__global__ void __launch_bounds__(256) mix_colscale_kernel(PrepArgs a) {
  __shared__ double red[256];
  const int k = blockIdx.x;
  if (k == 0 && threadIdx.x == 0) *a.err = 0;
  const double* W = a.W[k];
  double m = 0.0;
  for (int l = threadIdx.x; l < a.LW; l += 256) m = fmax(m, fabs(what_m(W, l, a.LW)));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    int e = 0;
    const double mm = red[0];
    (void)frexp(mm, &e);
    a.colscale[k] = mm > 0.0 ? ldexp(1.0, e) : 1.0;
  }
}

// one thread per l3: env(l3) and the column digits / scaled values
__global__ void __launch_bounds__(256) mix_bprep_kernel(PrepArgs a, int mode) {
  const int l3 = blockIdx.x * 256 + threadIdx.x;
  if (l3 >= a.envlen) return;
  double r = 0.0;
  for (int k = 0; k < a.K; ++k) r = fmax(r, fabs(what_m(a.W[k], l3, a.LW)) / a.colscale[k]);
  int e = 0;
  (void)frexp(r, &e);
  const double env = r > 0.0 ? ldexp(1.0, e) : 0.0;
  a.env[l3] = env;
  const int p = l3 & 1, h = l3 >> 1;
  for (int k = 0; k < a.kpad; ++k) {
    double b = 0.0;
    if (k < a.K && env > 0.0) b = what_m(a.W[k], l3, a.LW) / (a.colscale[k] * env);
    if (mode == MIX_I8) {
      // balanced base-256 digits of the integer nearest b 2^54 (|b| <= 1)
      long long nb = llrint(ldexp(b, kFixB));
      for (int t = 0; t < kNdB; ++t) {
        const long long d = ((nb + 128) & 255) - 128;
        nb = (nb - d) >> 8;
        const size_t off = ((((size_t)t * 2 + p) * a.hblocks + (h >> 4)) * a.kpad + k) * 16 +
                           (h & 15);
        a.bdig[off] = (int8_t)d;
      }
    } else {
      a.bf64[((size_t)p * (a.hblocks * 16) + h) * a.kpad + k] = b;
    }
  }
}

// ---- the contraction -------------------------------------------------------
// One wave: anti-diagonal S, rows x0 .. x0 + 16 RG - 1 (RG row groups of 16),
// all NCG column groups of 16.  Lane (j, g) = (lane % 16, lane / 16).
template <int MODE, int FAM, int RG, int NCG, int EPI, int COVK>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) mix_kernel(MixArgs A) {
  constexpr int kOdd = FAM == FAM_PO ? 1 : 0;
  constexpr int kUw = 16 * (RG - 1) + 80;  // staged window of each lane factor
  // per wave: lane-factor windows u1, u2 and the row-independent factor v
  __shared__ double lds_u[4][2][kUw];
  __shared__ double lds_v[4][64];
  __shared__ double lds_x[4][16][17];
  // B digits of one K-step, shared by the workgroup's four waves (their
  // anti-diagonals have one parity, so the same h rows): [column group]
  // [digit][lane], two buffers (the next step lands while this one is used)
  constexpr int kBcg = MODE == MIX_I8 ? NCG : 1, kBt = MODE == MIX_I8 ? kNdB : 1;
  __shared__ v4i lds_b0[kBcg][kBt][64];
  __shared__ v4i lds_b1[kBcg][kBt][64];

  const int lane = threadIdx.x & 63;
  const int wave = ufl(threadIdx.x >> 6);
  // tiles come in whole workgroups (ntiles is a multiple of 4): no wave
  // leaves before the barriers of the K-loop
  const int tile = ufl(blockIdx.x * 4 + wave);
  if (ufl(blockIdx.x * 4) >= A.ntiles) return;
  const int j = lane & 15, g = lane >> 4;
  const int S = ufl(A.tiles[tile].S);
  const int x0 = ufl(A.tiles[tile].x0);
  const int own_lo = ufl(A.tiles[tile].lo);
  const int own_hi = ufl(A.tiles[tile].hi);
  const int hb0 = ufl(A.tiles[tile].hb0);
  const int nsteps = ufl(A.tiles[tile].nsteps);

  // l3 = 2h + p = S - 2c - kOdd,  c = H - h
  const int p = (S - kOdd) & 1;
  const int H = (S - kOdd - p) >> 1;

  const double* T = kOdd ? A.Z : A.G;           // lane factors (index - kOdd)
  const double* Tc = kOdd ? A.Z : A.G;          // gamma_c / z_c
  const double* Ts = kOdd ? A.IG : A.D;         // delta_{S-c} / igamma_{S-c}
  double* su1 = lds_u[wave][0];
  double* su2 = lds_u[wave][1];
  double* sv = lds_v[wave];

  // spin-2 polynomial constants per (row group) of this lane's row j
  double sig[RG], kap[RG];
#pragma unroll
  for (int rg = 0; rg < RG; ++rg) {
    const double L1 = (double)(x0 + 16 * rg + j), L2 = (double)(S - (x0 + 16 * rg + j));
    const double P1 = L1 * (L1 + 1.0), P2 = L2 * (L2 + 1.0);
    sig[rg] = P1 + P2 - 1.0;
    kap[rg] = 1.0 + 2.0 * P1 * P2;
  }

  // The factors of one K-step (h = hb .. hb + 63): the lane-factor windows
  // u1[q] = T[x0 - H + hb + q - kOdd], u2[q] = T[S - x0 - 16 RG + 1 - H + hb + q - kOdd]
  // and the row-independent v[q] = Tc[c] Ts[S - c] env[l3] (h = hb + q).
  // They are loaded one step ahead into registers, untouched until the step
  // starts (so that no wait for them falls inside the current step), then
  // multiplied and written to this wave's LDS window.
  struct Fetch {
    double u1a, u1b, u2a, u2b, tc, ts, ev;
  };
  auto fetch = [&](int hb) {
    Fetch f;
    const int b1 = x0 - H + hb - kOdd + kTabOff;
    const int b2 = S - x0 - 16 * RG + 1 - H + hb - kOdd + kTabOff;
    f.u1a = T[b1 + lane];
    f.u2a = T[b2 + lane];
    f.u1b = lane + 64 < kUw ? T[b1 + lane + 64] : 0.0;
    f.u2b = lane + 64 < kUw ? T[b2 + lane + 64] : 0.0;
    const int h = hb + lane, c = H - h;
    f.tc = Tc[kTabOff + c];
    f.ts = Ts[kTabOff + S - c];
    f.ev = A.env[2 * h + p];
    return f;
  };
  auto commit = [&](const Fetch& f) {
    su1[lane] = f.u1a;
    su2[lane] = f.u2a;
    if (lane + 64 < kUw) {
      su1[lane + 64] = f.u1b;
      su2[lane + 64] = f.u2b;
    }
    sv[lane] = (f.tc * f.ts) * f.ev;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  auto unstage = [&]() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  // B digits of the K-step at hb into one LDS buffer, straight from global
  // memory (lane l's 16 bytes land at buf[cg][t][l]: lane (j, g) carries
  // h = hb + 16 g .. + 15 of column 16 cg + j); the workgroup's waves share
  // the NCG kNdB loads of a step
  auto issue_b = [&](int hb, v4i (*buf)[kBt][64]) {
    for (int n = wave; n < NCG * kNdB; n += 4) {
      const int cg = n / kNdB, t = n - cg * kNdB;
      __builtin_amdgcn_global_load_lds(
          (const __attribute__((address_space(1))) void*)(A.bdig +
              (((size_t)t * 2 + p) * A.hblocks + (hb >> 4) + g) * A.kpad + 16 * cg + j),
          (__attribute__((address_space(3))) void*)&buf[cg][t][0], 16, 0, 0);
    }
  };

  // ---- row scales (MIX_I8): the exact maximum of |A'| over each row's
  // family, from the row-maximum pass; rows outside l1 <= l2 in [lmin, lmax]
  // are computed but never written, their scale does not matter ----
  double rscale[RG];
  int rexp[RG];
  Fetch nf{};
#pragma unroll
  for (int rg = 0; rg < RG; ++rg) {
    rscale[rg] = 1.0;
    rexp[rg] = 0;
    if constexpr (MODE == MIX_I8) {
      const int l1 = x0 + 16 * rg + j, l2 = S - l1;
      double m = 0.0;
      if (l1 >= A.lmin && l1 <= l2 && l2 <= A.lmax)
        m = A.rowmax[A.doff[S - 2 * A.lmin] + (l1 - max(A.lmin, S - A.lmax))];
      int e = 0;
      (void)frexp(m, &e);
      rexp[rg] = m > 0.0 ? e : 0;
      rscale[rg] = ldexp(1.0, kFix - rexp[rg]);
    }
  }

  // ---- pass 2: the products ----
  // MIX_I8: acc[rg][cg][d] = sum over digit pairs with s + t = d + kDmin
  // MIX_F64: acc[rg][cg][0] in FP64
  v4i iacc[MODE == MIX_I8 ? RG : 1][MODE == MIX_I8 ? NCG : 1][kNdiag];
  v4d dacc[MODE == MIX_F64 ? RG : 1][MODE == MIX_F64 ? NCG : 1];
  if constexpr (MODE == MIX_I8) {
#pragma unroll
    for (int rg = 0; rg < RG; ++rg)
#pragma unroll
      for (int cg = 0; cg < NCG; ++cg)
#pragma unroll
        for (int d = 0; d < kNdiag; ++d) iacc[rg][cg][d] = v4i{0, 0, 0, 0};
  } else {
#pragma unroll
    for (int rg = 0; rg < RG; ++rg)
#pragma unroll
      for (int cg = 0; cg < NCG; ++cg) dacc[rg][cg] = v4d{0.0, 0.0, 0.0, 0.0};
  }

  if constexpr (MODE == MIX_I8) {
    // One K-step: every load of step st was issued a whole step earlier.
    auto step = [&](int st, v4i (*cur)[kBt][64], v4i (*nxt)[kBt][64]) {
      const int hb = hb0 + 64 * st;
      // this wave's loads for step st (its share of cur, its factors) have
      // landed; after the barrier every wave's share of cur has, and no wave
      // still reads nxt (its step st - 1 buffer)
      __builtin_amdgcn_s_waitcnt(0xF70);  // vmcnt(0)
      __syncthreads();
      commit(nf);
      if (st + 1 < nsteps) {
        issue_b(hb + 64, nxt);
        nf = fetch(hb + 64);
      }
      v4i ad[RG][kNd];
#pragma unroll
      for (int rg = 0; rg < RG; ++rg) {
        uint32_t w[4][kNd];
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) {
          uint32_t lo[4], hi[4];
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            const int q = 16 * g + 4 * qq + b;
            const double a =
                fam_value<FAM>(su1[16 * rg + j + q], su2[16 * (RG - 1 - rg) + 15 - j + q], sv[q],
                               sig[rg], kap[rg], 2 * (hb + q) + p);
            split_words(a * rscale[rg], lo[b], hi[b]);
          }
          transpose4(lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3], w[qq]);
        }
#pragma unroll
        for (int t = 0; t < kNd; ++t)
          ad[rg][t] = v4i{(int)w[0][t], (int)w[1][t], (int)w[2][t], (int)w[3][t]};
      }
#pragma unroll
      for (int cg = 0; cg < NCG; ++cg) {
        v4i bd[kNdB];
#pragma unroll
        for (int t = 0; t < kNdB; ++t) bd[t] = cur[cg][t][lane];
#pragma unroll
        for (int rg = 0; rg < RG; ++rg)
#pragma unroll
          for (int s2 = 0; s2 < kNd; ++s2)
#pragma unroll
            for (int t = 0; t < kNdB; ++t) {
              if (s2 + t < kDmin) continue;
              iacc[rg][cg][s2 + t - kDmin] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                  ad[rg][s2], bd[t], iacc[rg][cg][s2 + t - kDmin], 0, 0, 0);
            }
      }
      unstage();
    };
    if (nsteps > 0) {
      issue_b(hb0, lds_b0);
      nf = fetch(hb0);
    }
    for (int st = 0; st < nsteps; st += 2) {
      step(st, lds_b0, lds_b1);
      if (st + 1 < nsteps) step(st + 1, lds_b1, lds_b0);
    }
  } else {
    if (nsteps > 0) nf = fetch(hb0);
    for (int st = 0; st < nsteps; ++st) {
      const int hb = hb0 + 64 * st;
      commit(nf);
      if (st + 1 < nsteps) nf = fetch(hb + 64);
      // FP64: 16 MFMA steps of 4 c; lane (j, g) supplies h = hb + 4 m + g
#pragma unroll 4
      for (int m = 0; m < 16; ++m) {
        const int q = 4 * m + g;
        double bv[NCG];
#pragma unroll
        for (int cg = 0; cg < NCG; ++cg)
          bv[cg] = A.bf64[((size_t)p * (A.hblocks * 16) + hb + q) * A.kpad + 16 * cg + j];
#pragma unroll
        for (int rg = 0; rg < RG; ++rg) {
          const double a =
              fam_value<FAM>(su1[16 * rg + j + q], su2[16 * (RG - 1 - rg) + 15 - j + q], sv[q],
                             sig[rg], kap[rg], 2 * (hb + q) + p);
#pragma unroll
          for (int cg = 0; cg < NCG; ++cg)
            dacc[rg][cg] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv[cg], dacc[rg][cg], 0, 0, 0);
        }
      }
      unstage();
    }
  }
  if (own_hi < own_lo) return;

  // ---- epilogue: lane (j, g) holds four rows of column 16 cg + j: rows
  // 4 g + i in register i for the int8 MFMA, rows 4 i + g for the FP64 one ----
  const int lmin = A.lmin, lmax = A.lmax;
  bool bad = false;
  double* sx = &lds_x[wave][0][0];
#pragma unroll
  for (int rg = 0; rg < RG; ++rg) {
    int er[4], ri[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      ri[i] = MODE == MIX_I8 ? 4 * g + i : 4 * i + g;
      er[i] = MODE == MIX_I8 ? __shfl(rexp[rg], ri[i]) : 0;
    }
#pragma unroll
    for (int cg = 0; cg < NCG; ++cg) {
      const int k = 16 * cg + j;
      const double ck = k < A.ncol ? A.colscale[k] : 0.0;
      double xv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        double x;
        if constexpr (MODE == MIX_I8) {
          const v4i* c = iacc[rg][cg];
          x = (double)c[kNdiag - 1][i];
#pragma unroll
          for (int d = kNdiag - 2; d >= 0; --d) x = fma(x, 256.0, (double)c[d][i]);
          // sum A' B' = X 256^kDmin 2^(e - kFix - kFixB);  Xi = that times c_k
          x = ldexp(x, 8 * kDmin + er[i] - kFix - kFixB) * ck;
        } else {
          x = dacc[rg][cg][i] * ck;
        }
        xv[i] = x;
      }
      if constexpr (EPI == EPI_MCM) {
        // kappa Xi into the stage: for one i, the 4 g-rows of 16 consecutive
        // columns are 4 whole 128-byte lines
        const int64_t drow = A.doff[S - 2 * lmin] - max(lmin, S - lmax);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int l1 = x0 + 16 * rg + ri[i];
          const int l2 = S - l1;
          if (l1 < own_lo || l1 > own_hi || l1 < lmin || l1 > l2 || l2 > lmax) continue;
          const double L1 = (double)l1, L2 = (double)l2;
          double kk = 1.0;
          if constexpr (FAM != FAM_00) {
            const double pi1 = (L1 - 1.0) * L1 * (L1 + 1.0) * (L1 + 2.0);
            const double pi2 = (L2 - 1.0) * L2 * (L2 + 1.0) * (L2 + 2.0);
            const bool spin2 = (l1 >= 2) && (l2 >= 2);
            if constexpr (FAM == FAM_02) kk = spin2 ? 1.0 / (2.0 * sqrt(pi1 * pi2)) : 0.0;
            else kk = spin2 ? 1.0 / (4.0 * pi1 * pi2) : 0.0;
          }
          A.dmat[(drow + l1) * A.kpad + k] = kk * xv[i];
        }
      } else {
        // covariance: rows x columns of this column group through LDS, then
        // one (row, block) per lane
#pragma unroll
        for (int i = 0; i < 4; ++i) sx[ri[i] * 17 + j] = xv[i];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        constexpr int kNW = COVK == PSX_COV_TTTT || COVK == PSX_COV_EEEE ? 8
                            : COVK == PSX_COV_TTEE                       ? 2
                                                                         : 4;
        constexpr int kBpc = 16 / kNW;  // blocks per column group
        for (int pr = lane; pr < 16 * kBpc; pr += 64) {
          const int r = pr & 15, bb = pr >> 4;
          const int blk = (16 * cg) / kNW + bb;
          const int l1 = x0 + 16 * rg + r;
          const int l2 = S - l1;
          if (blk * kNW >= A.ncol || l1 < own_lo || l1 > own_hi || l1 < lmin || l1 > l2 ||
              l2 > lmax)
            continue;
          const double L1 = (double)l1, L2 = (double)l2;
          double kk = 1.0;
          if constexpr (FAM != FAM_00) {
            const double pi1 = (L1 - 1.0) * L1 * (L1 + 1.0) * (L1 + 2.0);
            const double pi2 = (L2 - 1.0) * L2 * (L2 + 1.0) * (L2 + 2.0);
            const bool spin2 = (l1 >= 2) && (l2 >= 2);
            if constexpr (FAM == FAM_02) kk = spin2 ? 1.0 / (2.0 * sqrt(pi1 * pi2)) : 0.0;
            else kk = spin2 ? 1.0 / (4.0 * pi1 * pi2) : 0.0;
          }
          double xi[kNW];
#pragma unroll
          for (int w = 0; w < kNW; ++w) xi[w] = kk * sx[r * 17 + bb * kNW + w];
          const double v = cov_combine<COVK>(l1, l2, xi, A.spec + 4 * blk, A.ratio + 4 * blk, bad);
          double* o = A.out[blk];
          o[(int64_t)(l1 - lmin) + (int64_t)(l2 - lmin) * A.ldo] = v;
          o[(int64_t)(l2 - lmin) + (int64_t)(l1 - lmin) * A.ldo] = v;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    }
  }
  if constexpr (EPI == EPI_COV) {
    if (bad) atomicOr(A.err, 1);
  }
}

// ---- stage -> outputs ---------------------------------------------------------
// One workgroup per (16 x 16 tile of (l1 <= l2) rows and columns, 16
// spectra): the tile's values (16 consecutive doubles per pair in the stage)
// go through LDS, then both triangles of each output are written as column
// segments of 16 rows: M_k[l1, l2] = (2 l2 + 1) v, M_k[l2, l1] = (2 l1 + 1) v.
// Tiles come in 4 x 4 groups (from a host list), so that neighbouring
// workgroups write neighbouring segments of both triangles.
struct ScatterArgs {
  const double* dmat;
  const int64_t* doff;
  const int2* tiles;   // (ta, tb), ta <= tb, in 16-row units
  double* const* out;
  int64_t ldo;
  int lmin, lmax, kpad, ncol;
};

__global__ void __launch_bounds__(256) mix_scatter_kernel(ScatterArgs A) {
  __shared__ double buf[16][16][17];   // [k][a][b]
  __shared__ int64_t srow[31];         // stage row of (l1 = 0) on each anti-diagonal
  const int2 tt = A.tiles[blockIdx.x];
  const int kc = (int)blockIdx.y;
  const int a0 = A.lmin + 16 * tt.x, b0 = A.lmin + 16 * tt.y;
  const int tid = (int)threadIdx.x;
  if (tid < 31) {
    const int S = a0 + b0 + tid;
    srow[tid] = S <= 2 * A.lmax ? A.doff[S - 2 * A.lmin] - max(A.lmin, S - A.lmax) : 0;
  }
  __syncthreads();
  // load: 8 threads per (a, b) pair, 16 bytes each
#pragma unroll
  for (int pass = 0; pass < 8; ++pass) {
    const int pr = (tid >> 3) + 32 * pass, part = tid & 7;
    const int al = pr >> 4, bl = pr & 15;
    const int a = a0 + al, b = b0 + bl;
    double2 v = make_double2(0.0, 0.0);
    if (a <= b && b <= A.lmax)
      v = *(const double2*)(A.dmat + (srow[al + bl] + a) * A.kpad + 16 * kc + 2 * part);
    buf[2 * part][al][bl] = v.x;
    buf[2 * part + 1][al][bl] = v.y;
  }
  __syncthreads();
  // direct elements, rows a consecutive; then the mirror, rows b consecutive
#pragma unroll 4
  for (int pass = 0; pass < 16; ++pass) {
    const int idx = tid + 256 * pass;
    const int kl = idx >> 8, cl = (idx >> 4) & 15, rl = idx & 15;
    const int k = 16 * kc + kl;
    if (k >= A.ncol) continue;
    double* o = A.out[k];
    {
      const int a = a0 + rl, b = b0 + cl;
      if (a <= b && b <= A.lmax)
        o[(int64_t)(a - A.lmin) + (int64_t)(b - A.lmin) * A.ldo] =
            (2.0 * (double)b + 1.0) * buf[kl][rl][cl];
    }
    {
      const int a = a0 + cl, b = b0 + rl;
      if (a <= b && b <= A.lmax)
        o[(int64_t)(b - A.lmin) + (int64_t)(a - A.lmin) * A.ldo] =
            (2.0 * (double)a + 1.0) * buf[kl][cl][rl];
    }
  }
}

// The scatter's tile list: (ta <= tb) tiles of an nt x nt grid, in 4 x 4
// groups along the upper triangle.
#ifndef PSX_SCATTER_GROUP
#define PSX_SCATTER_GROUP 4
#endif
std::vector<int2> make_scatter_tiles(int nt) {
  constexpr int G = PSX_SCATTER_GROUP;
  std::vector<int2> out;
  out.reserve((size_t)nt * (nt + 1) / 2);
  for (int sb = 0; sb < nt; sb += G)
    for (int sa = 0; sa <= sb; sa += G)
      for (int tb = sb; tb < std::min(nt, sb + G); ++tb)
        for (int ta = sa; ta < std::min(nt, sa + G); ++ta)
          if (ta <= tb) out.push_back(make_int2(ta, tb));
  return out;
}

// ---- tiles -------------------------------------------------------------------
// For each anti-diagonal S the owned rows l1 in [max(lmin, S - lmax), min(S/2, lmax)]
// are cut into chunks of `rows` from the diagonal down.  A workgroup takes the
// same chunk of 4 anti-diagonals of one parity (S, S + 2, S + 4, S + 6): their
// B rows are the same for the same h, so the four waves share one B stream
// over the union of their K-ranges (the family products vanish outside a
// wave's own range: zero table padding below it, zero What above LW - 1).
// A short group is filled with inactive copies (no owned rows) of its first
// tile.  Workgroups longest first, laid out in runs per XCD as the VALU
// kernels' tiles.
std::vector<MTile> make_mix_tiles(int lmin, int lmax, int LW, int odd, int rows) {
  std::vector<std::vector<MTile>> chunks[2];  // [parity][chunk] -> tiles by S
  for (int S = 2 * lmin; S <= 2 * lmax; ++S) {
    const int hi = std::min(S / 2, lmax);
    const int lo = std::max(lmin, S - lmax);
    const int p = (S - odd) & 1, H = (S - odd - p) >> 1;
    const int clo = std::max(0, (S - odd - (LW - 1) + 1) >> 1);
    const int hend = H - clo;
    int k = 0;
    for (int top = hi; top >= lo; top -= rows, ++k) {
      MTile t;
      t.S = S;
      t.x0 = std::max(top - rows + 1, 0);
      t.lo = std::max(lo, top - rows + 1);
      t.hi = top;
      // own K-range [hb0, hend] (nsteps = 0: nothing to contract)
      t.hb0 = std::max(0, H - (t.x0 + rows - 1)) & ~15;
      t.nsteps = hend;  // last h, replaced by the group's step count below
      auto& ck = chunks[p];
      if ((int)ck.size() <= k) ck.resize(k + 1);
      ck[k].push_back(t);
    }
  }
  struct WG {
    MTile t[4];
    int cost;
  };
  std::vector<WG> wgs;
  for (auto& par : chunks)
    for (auto& ck : par)
      for (size_t i = 0; i < ck.size(); i += 4) {
        WG g;
        const size_t n = std::min(ck.size(), i + 4) - i;
        int hb0 = 1 << 30, hend = -1;
        for (size_t q = 0; q < n; ++q) {
          hb0 = std::min(hb0, ck[i + q].hb0);
          hend = std::max(hend, ck[i + q].nsteps);
        }
        const int nsteps = hend >= hb0 ? (hend - hb0 + 64) >> 6 : 0;
        for (size_t q = 0; q < 4; ++q) {
          g.t[q] = ck[i + (q < n ? q : 0)];
          if (q >= n) {
            g.t[q].lo = 1;
            g.t[q].hi = 0;
          }
          g.t[q].hb0 = nsteps > 0 ? hb0 : 0;
          g.t[q].nsteps = nsteps;
        }
        g.cost = nsteps;
        wgs.push_back(g);
      }
  int maxc = 1;
  for (const WG& g : wgs) maxc = std::max(maxc, g.cost);
  const int bucket = std::max(1, maxc / 32);
  std::stable_sort(wgs.begin(), wgs.end(),
                   [bucket](const WG& a, const WG& b) { return a.cost / bucket > b.cost / bucket; });
  constexpr int kXcds = 8, kRun = 8;
  const int nwg = (int)wgs.size();
  std::vector<MTile> out;
  out.reserve((size_t)nwg * 4);
  for (int b = 0; (int)out.size() < nwg * 4 && b <= 4 * nwg + 64 * kXcds * kRun; ++b) {
    const int x = b % kXcds, m = b / kXcds;
    const long w = (long)(m / kRun) * kXcds * kRun + (long)x * kRun + m % kRun;
    if (w >= nwg) continue;
    for (int q = 0; q < 4; ++q) out.push_back(wgs[w].t[q]);
  }
  return out;
}

// Host check of every index a tile set can touch (the kernel's access list:
// staged lane factors T[l - c - odd] over the window, T[c], Ts[S - c],
// env[l3], B rows h over the workgroup's K-range) against the buffer sizes,
// and of the workgroup invariants the shared B stream relies on.
void check_mix_geometry(const std::vector<MTile>& tiles, int lmax, int odd, int RG, int ntab,
                        int hblocks, int envlen) {
  PSX_REQUIRE(tiles.size() % 4 == 0, PSX_ERR_ARGUMENT, "mixed tiles: not whole workgroups");
  for (size_t i = 0; i < tiles.size(); ++i) {
    const MTile& t = tiles[i];
    const MTile& t0 = tiles[i & ~size_t(3)];
    PSX_REQUIRE(t.hb0 == t0.hb0 && t.nsteps == t0.nsteps && ((t.S - t0.S) & 1) == 0 &&
                    t.hb0 >= 0 && (t.hb0 & 15) == 0 && t.nsteps >= 0,
                PSX_ERR_ARGUMENT, "mixed tile geometry: workgroup K-range");
    PSX_REQUIRE(t.x0 >= 0 && t.x0 + 16 * RG - 1 <= lmax + 16 * RG && t.S >= 0 && t.S <= 2 * lmax,
                PSX_ERR_ARGUMENT, "mixed tile geometry: bad tile");
    if (t.hi >= t.lo)
      PSX_REQUIRE(t.lo >= t.x0 && t.hi <= t.x0 + 16 * RG - 1 && t.hi <= lmax, PSX_ERR_ARGUMENT,
                  "mixed tile geometry: bad row range");
    if (t.nsteps == 0) continue;
    const int S = t.S, p = (S - odd) & 1, H = (S - odd - p) >> 1;
    const int hb0 = t.hb0;
    const int hlast = hb0 + 64 * t.nsteps - 1;
    const int kUw = 16 * (RG - 1) + 80;
    const long u1lo = (long)t.x0 - H + hb0 - odd, u1hi = (long)t.x0 - H + hlast - 63 - odd + kUw - 1;
    const long b2 = (long)S - t.x0 - 16 * RG + 1 - H - odd;
    const long u2lo = b2 + hb0, u2hi = b2 + hlast - 63 + kUw - 1;
    const long clo_i = (long)H - hlast, chi_i = (long)H - hb0;
    const long slo = (long)S - chi_i, shi = (long)S - clo_i;
    const long lo = std::min({u1lo, u2lo, clo_i, slo}), hi = std::max({u1hi, u2hi, chi_i, shi});
    PSX_REQUIRE(lo >= -kTabOff && hi + kTabOff < ntab, PSX_ERR_ARGUMENT,
                "mixed tile geometry: table index out of range");
    PSX_REQUIRE(2L * hlast + p < envlen && (hlast >> 4) < hblocks, PSX_ERR_ARGUMENT,
                "mixed tile geometry: column table out of range");
  }
}

struct MixTileKey {
  int lmin, lmax, LW, odd, rows;
  bool operator<(const MixTileKey& o) const {
    return std::tie(lmin, lmax, LW, odd, rows) < std::tie(o.lmin, o.lmax, o.LW, o.odd, o.rows);
  }
};
struct MixTiles {
  MTile* d = nullptr;
  int n = 0;
  // the output stage's row offsets: doff[S - 2 lmin] = pairs (l1 <= l2) on
  // the anti-diagonals before S; npairs in all
  int64_t* doff = nullptr;
  int64_t npairs = 0;
  int2* stiles = nullptr;   // the scatter's tile list
  int nstiles = 0;
};
std::map<std::pair<void*, MixTileKey>, MixTiles> g_mix_tiles;
std::mutex g_mix_mu;

using MixFn = void (*)(MixArgs);

template <int MODE, int FAM, int RG, int EPI, int COVK>
MixFn mix_fn(int ncg) {
  switch (ncg) {
    case 1: return mix_kernel<MODE, FAM, RG, 1, EPI, COVK>;
    case 2: return mix_kernel<MODE, FAM, RG, 2, EPI, COVK>;
    case 3: return mix_kernel<MODE, FAM, RG, 3, EPI, COVK>;
    default: return mix_kernel<MODE, FAM, RG, 4, EPI, COVK>;
  }
}

#ifndef PSX_MIX_RG
#define PSX_MIX_RG 1
#endif

template <int MODE>
MixFn pick_mix(int fam, int epi, int covk, int ncg) {
  constexpr int RG = PSX_MIX_RG;
  if (epi == EPI_MCM) {
    switch (fam) {
      case FAM_00: return mix_fn<MODE, FAM_00, RG, EPI_MCM, 0>(ncg);
      case FAM_02: return mix_fn<MODE, FAM_02, RG, EPI_MCM, 0>(ncg);
      case FAM_PE: return mix_fn<MODE, FAM_PE, RG, EPI_MCM, 0>(ncg);
      case FAM_PO: return mix_fn<MODE, FAM_PO, RG, EPI_MCM, 0>(ncg);
    }
  } else {
    switch (covk) {
      case PSX_COV_TTTT: return mix_fn<MODE, FAM_00, RG, EPI_COV, PSX_COV_TTTT>(ncg);
      case PSX_COV_EEEE: return mix_fn<MODE, FAM_PE, RG, EPI_COV, PSX_COV_EEEE>(ncg);
      case PSX_COV_TTTE: return mix_fn<MODE, FAM_00, RG, EPI_COV, PSX_COV_TTTE>(ncg);
      case PSX_COV_TEEE: return mix_fn<MODE, FAM_02, RG, EPI_COV, PSX_COV_TEEE>(ncg);
      case PSX_COV_TEEE_PLANCK:
        return mix_fn<MODE, FAM_PE, RG, EPI_COV, PSX_COV_TEEE_PLANCK>(ncg);
      case PSX_COV_TTEE: return mix_fn<MODE, FAM_00, RG, EPI_COV, PSX_COV_TTEE>(ncg);
    }
  }
  throw Error{PSX_ERR_ARGUMENT, "the matrix-core contraction does not take this kind"};
}

size_t al256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

void release_mix_tiles(Ctx* c) {
  std::lock_guard<std::mutex> lk(g_mix_mu);
  for (auto it = g_mix_tiles.begin(); it != g_mix_tiles.end();) {
    if (it->first.first == (void*)c)
      it = g_mix_tiles.erase(it);
    else
      ++it;
  }
}

int mix_family_of_cov(int kind) {
  switch (kind) {
    case PSX_COV_TTTT: case PSX_COV_TTTE: case PSX_COV_TTEE: return FAM_00;
    case PSX_COV_EEEE: case PSX_COV_TEEE_PLANCK: return FAM_PE;
    case PSX_COV_TEEE: return FAM_02;
  }
  throw Error{PSX_ERR_ARGUMENT,
              "the matrix-core contraction takes TTTT, EEEE, TTTE, TEEE, TEEE-Planck and TTEE "
              "(TETE mixes two families)"};
}

// One pass of the matrix-core contraction on the context stream.
void run_mixed(Ctx* c, const MixRequest& rq) {
  const int K = rq.ncol;
  PSX_REQUIRE(K >= 1 && K <= kMaxCols, PSX_ERR_ARGUMENT, "between 1 and 64 W spectra per pass");
  PSX_REQUIRE(rq.lmin >= 0 && rq.lmax >= rq.lmin && rq.lmax <= 8191, PSX_ERR_ARGUMENT,
              "need 0 <= lmin <= lmax <= 8191");
  PSX_REQUIRE(rq.LW >= 1 && rq.LW < (1 << 24), PSX_ERR_ARGUMENT, "bad W length");
  PSX_REQUIRE(rq.mode == PSX_MIXED || rq.mode == PSX_FP64_MFMA, PSX_ERR_ARGUMENT,
              "matrix-core precision must be PSX_MIXED or PSX_FP64_MFMA");
  const int mode = rq.mode == PSX_MIXED ? MIX_I8 : MIX_F64;
  const int LW = (int)std::min<int64_t>(rq.LW, 2 * (int64_t)rq.lmax + 1);
  const int n = rq.lmax - rq.lmin + 1;
  PSX_REQUIRE(rq.ldo >= n, PSX_ERR_AXES, "leading dimension smaller than lmax-lmin+1");
  const int nout = rq.epi == EPI_MCM ? K : rq.nblk;
  if (rq.epi == EPI_COV) {
    PSX_REQUIRE(rq.nblk >= 1 && rq.nblk <= kMaxBlk && rq.nblk * rq.nW == K, PSX_ERR_ARGUMENT,
                "covariance batch: nblk x nW columns, at most 16 blocks");
    PSX_REQUIRE(mix_family_of_cov(rq.cov_kind) == rq.fam, PSX_ERR_ARGUMENT,
                "covariance kind and family disagree");
  }
  const int kpad = (K + 15) & ~15, ncg = kpad / 16;
  const int odd = rq.fam == FAM_PO ? 1 : 0;
  constexpr int RG = PSX_MIX_RG;
  const int hblocks = (rq.lmax + 64 + 15) / 16 + 2;
  const int envlen = 32 * hblocks;
  ensure_tables(c, LW + 2 * rq.lmax + 512);

  const MixTiles* mt;
  {
    std::lock_guard<std::mutex> lk(g_mix_mu);
    auto key = std::make_pair((void*)c, MixTileKey{rq.lmin, rq.lmax, LW, odd, 16 * RG});
    auto it = g_mix_tiles.find(key);
    if (it == g_mix_tiles.end()) {
      std::vector<MTile> h = make_mix_tiles(rq.lmin, rq.lmax, LW, odd, 16 * RG);
      check_mix_geometry(h, rq.lmax, odd, RG, table_entries(LW + 2 * rq.lmax + 512), hblocks,
                         envlen);
      MixTiles m;
      m.n = (int)h.size();
      m.d = (MTile*)ctx_alloc(c, sizeof(MTile) * std::max<size_t>(1, h.size()));
      PSX_HIP(hipMemcpy(m.d, h.data(), sizeof(MTile) * h.size(), hipMemcpyHostToDevice));
      std::vector<int64_t> off(2 * (size_t)(rq.lmax - rq.lmin) + 1);
      int64_t acc = 0;
      for (int S = 2 * rq.lmin; S <= 2 * rq.lmax; ++S) {
        off[S - 2 * rq.lmin] = acc;
        acc += std::max(0, std::min(S / 2, rq.lmax) - std::max(rq.lmin, S - rq.lmax) + 1);
      }
      m.npairs = acc;
      m.doff = (int64_t*)ctx_alloc(c, sizeof(int64_t) * off.size());
      PSX_HIP(hipMemcpy(m.doff, off.data(), sizeof(int64_t) * off.size(), hipMemcpyHostToDevice));
      const std::vector<int2> st = make_scatter_tiles((rq.lmax - rq.lmin + 1 + 15) / 16);
      m.nstiles = (int)st.size();
      m.stiles = (int2*)ctx_alloc(c, sizeof(int2) * st.size());
      PSX_HIP(hipMemcpy(m.stiles, st.data(), sizeof(int2) * st.size(), hipMemcpyHostToDevice));
      it = g_mix_tiles.emplace(key, m).first;
    }
    mt = &it->second;
  }
  PSX_REQUIRE(c->tables.n >= table_entries(LW + 2 * rq.lmax + 512), PSX_ERR_ARGUMENT,
              "tables smaller than the mixed geometry needs");

  // scratch: [colscale][env][B table][pointer tables][flag]
  const size_t b_cs = al256(sizeof(double) * kpad), b_env = al256(sizeof(double) * envlen);
  const size_t b_tab = mode == MIX_I8 ? al256((size_t)kNdB * 2 * hblocks * kpad * 16)
                                      : al256(sizeof(double) * 2 * (size_t)hblocks * 16 * kpad);
  const size_t b_ptr = al256(sizeof(void*) * (kMaxCols + 2 * 4 * kMaxBlk + kMaxCols));
  // EPI_MCM: the anti-diagonal-major output stage, npairs x kpad doubles
  const size_t b_stage = rq.epi == EPI_MCM ? al256(sizeof(double) * (size_t)mt->npairs * kpad) : 0;
  // MIX_I8: the row maxima (npairs doubles) and the row-maximum pass's records
  ContractRequest rr;
  rr.lmin = rq.lmin;
  rr.lmax = rq.lmax;
  rr.LW = LW;
  rr.nfam[rq.fam] = 1;
  rr.epi = EPI_RMAX;
  rr.ldo = n;
  rr.doff = mt->doff;
  const size_t b_rmax = mode == MIX_I8 ? al256(sizeof(double) * (size_t)mt->npairs) : 0;
  const size_t b_rrec = mode == MIX_I8 ? al256(contraction_scratch_bytes(rr)) : 0;
  const size_t need = b_cs + b_env + b_tab + b_ptr + 256 + b_stage + b_rmax + b_rrec;
  char* base = (char*)ctx_scratch(c, need);
  double* colscale = (double*)base;
  double* env = (double*)(base + b_cs);
  char* tab = base + b_cs + b_env;
  void** ptrs = (void**)(base + b_cs + b_env + b_tab);
  int* err = (int*)(base + b_cs + b_env + b_tab + b_ptr);
  char* tail = base + b_cs + b_env + b_tab + b_ptr + 256;
  double* stage = rq.epi == EPI_MCM ? (double*)tail : nullptr;
  double* rowmax = mode == MIX_I8 ? (double*)(tail + b_stage) : nullptr;
  void* rrec = mode == MIX_I8 ? (void*)(tail + b_stage + b_rmax) : nullptr;

  // pointer tables: W sources, outputs, spectra, ratios (one copy)
  std::vector<const void*> hp(kMaxCols + 2 * 4 * kMaxBlk + kMaxCols, nullptr);
  for (int k = 0; k < K; ++k) hp[k] = rq.dW[k];
  for (int k = 0; k < nout; ++k) hp[kMaxCols + k] = rq.dout[k];
  if (rq.epi == EPI_COV)
    for (int b = 0; b < rq.nblk; ++b)
      for (int i = 0; i < 4; ++i) {
        hp[2 * kMaxCols + 4 * b + i] = rq.dspec[b][i];
        hp[2 * kMaxCols + 4 * kMaxBlk + 4 * b + i] = rq.dratio[b][i];
      }
  PSX_HIP(hipMemcpyAsync(ptrs, hp.data(), sizeof(void*) * hp.size(), hipMemcpyHostToDevice,
                         c->stream));

  PrepArgs pa{};
  pa.W = (const double* const*)ptrs;
  pa.K = K;
  pa.kpad = kpad;
  pa.LW = LW;
  pa.colscale = colscale;
  pa.env = env;
  pa.envlen = envlen;
  pa.bdig = mode == MIX_I8 ? (int8_t*)tab : nullptr;
  pa.bf64 = mode == MIX_F64 ? (double*)tab : nullptr;
  pa.hblocks = hblocks;
  pa.err = err;
  {
    TimedLaunch tl(c, "psx_mix_prep");
    hipLaunchKernelGGL(mix_colscale_kernel, dim3(K), dim3(256), 0, c->stream, pa);
    hipLaunchKernelGGL(mix_bprep_kernel, dim3((envlen + 255) / 256), dim3(256), 0, c->stream, pa,
                       mode);
  }

  if (mode == MIX_I8) {
    // row maxima of |A'| = |family product| env(l3) over l3 <= LW - 1
    rr.dW[0] = env;
    rr.dout[0] = rowmax;
    (void)run_contraction_on(c, rr, c->stream, rrec, b_rrec);
  }

  MixArgs ma{};
  ma.tiles = mt->d;
  ma.ntiles = mt->n;
  ma.lmin = rq.lmin;
  ma.lmax = rq.lmax;
  ma.LW = LW;
  ma.G = c->tables.dG;
  ma.Z = c->tables.dZ;
  ma.D = c->tables.dD;
  ma.IG = c->tables.dIG;
  ma.env = env;
  ma.bdig = mode == MIX_I8 ? (const v4i*)tab : nullptr;
  ma.bf64 = mode == MIX_F64 ? (const double*)tab : nullptr;
  ma.hblocks = hblocks;
  ma.kpad = kpad;
  ma.ncol = K;
  ma.colscale = colscale;
  ma.out = (double* const*)(ptrs + kMaxCols);
  ma.ldo = rq.ldo;
  ma.nW = rq.nW;
  ma.spec = (const double* const*)(ptrs + 2 * kMaxCols);
  ma.ratio = (const double* const*)(ptrs + 2 * kMaxCols + 4 * kMaxBlk);
  ma.err = err;
  ma.dmat = stage;
  ma.doff = mt->doff;
  ma.rowmax = rowmax;
  const MixFn fn = mode == MIX_I8 ? pick_mix<MIX_I8>(rq.fam, rq.epi, rq.cov_kind, ncg)
                                  : pick_mix<MIX_F64>(rq.fam, rq.epi, rq.cov_kind, ncg);
  {
    TimedLaunch tl(c, mode == MIX_I8 ? "psx_mix_i8" : "psx_mix_f64");
    hipLaunchKernelGGL(fn, dim3((unsigned)((mt->n + 3) / 4)), dim3(256), 0, c->stream, ma);
  }
  if (rq.epi == EPI_MCM) {
    ScatterArgs sa{};
    sa.dmat = stage;
    sa.doff = mt->doff;
    sa.out = (double* const*)(ptrs + kMaxCols);
    sa.ldo = rq.ldo;
    sa.lmin = rq.lmin;
    sa.lmax = rq.lmax;
    sa.kpad = kpad;
    sa.ncol = K;
    sa.tiles = mt->stiles;
    TimedLaunch tl(c, "psx_mix_scatter");
    hipLaunchKernelGGL(mix_scatter_kernel, dim3((unsigned)mt->nstiles, (unsigned)ncg), dim3(256),
                       0, c->stream, sa);
  }
  if (rq.epi == EPI_COV) {
    int h = 0;
    PSX_HIP(hipMemcpyAsync(&h, err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    PSX_HIP(hipStreamSynchronize(c->stream));
    PSX_REQUIRE(h == 0, PSX_ERR_DOMAIN, "sqrt of a negative spectrum product (Julia DomainError)");
  }
}

}  // namespace psx
