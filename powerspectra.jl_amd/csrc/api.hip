// extern "C" entry points for mode-coupling matrices and covariance blocks.
#include "psx_internal.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace psx {

int mcm_family(int kind) {
  switch (kind) {
    case PSX_MCM_00: return FAM_00;
    case PSX_MCM_02: return FAM_02;
    case PSX_MCM_PP: return FAM_PE;
    case PSX_MCM_MM: return FAM_PO;
  }
  throw Error{PSX_ERR_ARGUMENT, "unknown mcm kind (valid: 00, 02, PP, MM)"};
}

// W -> V table order for each covariance kind, and the family of each table
// (reference argument order W1..WK of loop_cov*!).
void cov_layout(int kind, ContractRequest& rq, const double* const* W) {
  int nW = 0;
  PSX_REQUIRE(psx_cov_arity(kind, &nW, nullptr, nullptr) == PSX_OK, PSX_ERR_ARGUMENT,
              "unknown covariance kind");
  auto put = [&](int fam, std::initializer_list<int> idx) {
    int k = rq.nfam[0] + rq.nfam[1] + rq.nfam[2] + rq.nfam[3];
    for (int i : idx) rq.dW[k++] = W[i];
    rq.nfam[fam] += (int)idx.size();
  };
  switch (kind) {
    case PSX_COV_TTTT: put(FAM_00, {0, 1, 2, 3, 4, 5, 6, 7}); break;
    case PSX_COV_EEEE: put(FAM_PE, {0, 1, 2, 3, 4, 5, 6, 7}); break;
    case PSX_COV_TTTE: put(FAM_00, {0, 1, 2, 3}); break;
    case PSX_COV_TETE:
      put(FAM_00, {1});
      put(FAM_02, {0, 2, 3, 4});
      break;
    case PSX_COV_TEEE: put(FAM_02, {0, 1, 2, 3}); break;
    case PSX_COV_TEEE_PLANCK: put(FAM_PE, {0, 1, 2, 3}); break;
    case PSX_COV_TTEE: put(FAM_00, {0, 1}); break;
  }
}

}  // namespace psx

using namespace psx;

namespace {

Ctx* C(psx_ctx* h) { return reinterpret_cast<Ctx*>(h); }

void check_block(int lmin, int lmax, int64_t LW, int64_t ld) {
  PSX_REQUIRE(lmin >= 0 && lmax >= lmin, PSX_ERR_ARGUMENT, "need 0 <= lmin <= lmax");
  PSX_REQUIRE(lmax <= 8191, PSX_ERR_ARGUMENT, "lmax above 8191 is not supported");
  PSX_REQUIRE(LW >= 1, PSX_ERR_ARGUMENT, "W must have at least one entry");
  PSX_REQUIRE(ld >= lmax - lmin + 1, PSX_ERR_AXES,
              "leading dimension smaller than lmax-lmin+1");
}

void h2d(Ctx* c, void* dst, const void* src, size_t bytes) {
  PSX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
}

ContractRequest block_request(int lmin, int lmax, int64_t LW, int64_t ld) {
  ContractRequest rq;
  rq.lmin = lmin;
  rq.lmax = lmax;
  rq.LW = LW;
  rq.ldo = ld;
  rq.epi = EPI_MCM;
  return rq;
}

ContractRequest mcm_request(int kind, int lmin, int lmax, const double* dW, int64_t LW,
                            double* dM, int64_t ldM) {
  ContractRequest rq = block_request(lmin, lmax, LW, ldM);
  rq.nfam[mcm_family(kind)] = 1;
  rq.dW[0] = dW;
  rq.dout[0] = dM;
  return rq;
}

ContractRequest pm_request(int lmin, int lmax, const double* dW, int64_t LW, double* dMpp,
                           double* dMmm, int64_t ldM) {
  ContractRequest rq = block_request(lmin, lmax, LW, ldM);
  rq.nfam[FAM_PE] = 1;
  rq.nfam[FAM_PO] = 1;
  rq.dW[0] = rq.dW[1] = dW;
  rq.dout[0] = dMpp;
  rq.dout[1] = dMmm;
  return rq;
}

// TT, TE, ET, ++ and -- from the four W spectra (TT, TE, ET, PP) in one launch
ContractRequest set_request(int lmin, int lmax, int64_t LW, const double* const dW[4],
                            double* const dM[5], int64_t ldM) {
  ContractRequest rq = block_request(lmin, lmax, LW, ldM);
  rq.nfam[FAM_00] = 1;
  rq.nfam[FAM_02] = 2;
  rq.nfam[FAM_PE] = 1;
  rq.nfam[FAM_PO] = 1;
  const double* Ws[5] = {dW[0], dW[1], dW[2], dW[3], dW[3]};
  for (int k = 0; k < 5; ++k) {
    rq.dW[k] = Ws[k];
    rq.dout[k] = dM[k];
  }
  return rq;
}

ContractRequest cov_request(int kind, int lmin, int lmax, const double* const* dspectra,
                            const double* const* dratios, const double* const* dW, int64_t LW,
                            int precision, double* dC, int64_t ldC) {
  int nW = 0, nS = 0, nR = 0;
  PSX_REQUIRE(psx_cov_arity(kind, &nW, &nS, &nR) == PSX_OK, PSX_ERR_ARGUMENT,
              "unknown covariance kind");
  PSX_REQUIRE(precision == PSX_FP64 || precision == PSX_MIXED || precision == PSX_FP64_MFMA,
              PSX_ERR_ARGUMENT, "precision must be PSX_FP64, PSX_MIXED or PSX_FP64_MFMA");
  PSX_REQUIRE(dW && dC && dspectra, PSX_ERR_ARGUMENT, "null pointer");
  for (int i = 0; i < nW; ++i) PSX_REQUIRE(dW[i], PSX_ERR_ARGUMENT, "null W spectrum");
  for (int i = 0; i < nS; ++i) PSX_REQUIRE(dspectra[i], PSX_ERR_ARGUMENT, "null spectrum");
  for (int i = 0; i < nR; ++i)
    PSX_REQUIRE(dratios && dratios[i], PSX_ERR_ARGUMENT, "null noise ratio");
  ContractRequest rq = block_request(lmin, lmax, LW, ldC);
  cov_layout(kind, rq, dW);
  rq.epi = EPI_COV;
  rq.cov_kind = kind;
  rq.dout[0] = dC;
  for (int i = 0; i < 4; ++i) {
    rq.dspec[i] = i < nS ? dspectra[i] : nullptr;
    rq.dratio[i] = i < nR ? dratios[i] : nullptr;
  }
  rq.precision = precision;
  return rq;
}

// column groups of 16 spectra per matrix-core pass (register budget of the
// int8 kernel: seven int32 accumulators per row group and column group)
#ifndef PSX_MIX_NCG_MAX
#define PSX_MIX_NCG_MAX 4
#endif

void check_precision(int precision) {
  PSX_REQUIRE(precision == PSX_FP64 || precision == PSX_MIXED || precision == PSX_FP64_MFMA,
              PSX_ERR_ARGUMENT, "precision must be PSX_FP64, PSX_MIXED or PSX_FP64_MFMA");
}

// K blocks of one mode-coupling kind on device buffers: FP64 on the vector
// ALUs in groups of 8, 4, 2, 1 spectra per launch, or the matrix-core pass.
void mcm_batch_dev(Ctx* c, int kind, int lmin, int lmax, int K, const double* const* dW,
                   int64_t LW, int precision, double* const* dM, int64_t ldM) {
  PSX_REQUIRE(K >= 1 && K <= 64, PSX_ERR_ARGUMENT, "between 1 and 64 W spectra");
  PSX_REQUIRE(dW && dM, PSX_ERR_ARGUMENT, "null pointer");
  for (int k = 0; k < K; ++k) PSX_REQUIRE(dW[k] && dM[k], PSX_ERR_ARGUMENT, "null pointer");
  const int fam = mcm_family(kind);
  if (precision == PSX_FP64) {
    for (int k0 = 0; k0 < K;) {
      const int left = K - k0;
      const int g = left >= 8 ? 8 : left >= 4 ? 4 : left >= 2 ? 2 : 1;
      ContractRequest rq = block_request(lmin, lmax, LW, ldM);
      rq.nfam[fam] = g;
      for (int i = 0; i < g; ++i) {
        rq.dW[i] = dW[k0 + i];
        rq.dout[i] = dM[k0 + i];
      }
      run_contraction(c, rq);
      k0 += g;
    }
    return;
  }
  // matrix-core passes of at most 16 PSX_MIX_NCG_MAX spectra each
  const int per = 16 * PSX_MIX_NCG_MAX;
  for (int k0 = 0; k0 < K; k0 += per) {
    const int nk = std::min(per, K - k0);
    MixRequest mr;
    mr.lmin = lmin;
    mr.lmax = lmax;
    mr.LW = LW;
    mr.fam = fam;
    mr.ncol = nk;
    for (int k = 0; k < nk; ++k) {
      mr.dW[k] = dW[k0 + k];
      mr.dout[k] = dM[k0 + k];
    }
    mr.epi = EPI_MCM;
    mr.ldo = ldM;
    mr.mode = precision;
    run_mixed(c, mr);
  }
}

// nblk covariance blocks of one kind on device buffers; the arrays hold the
// blocks' spectra, ratios and W spectra one block after another, each in the
// reference loop's argument order.
void cov_batch_dev(Ctx* c, int kind, int lmin, int lmax, int nblk,
                   const double* const* dspectra, const double* const* dratios,
                   const double* const* dW, int64_t LW, int precision, double* const* dC,
                   int64_t ldC) {
  int nW = 0, nS = 0, nR = 0;
  PSX_REQUIRE(psx_cov_arity(kind, &nW, &nS, &nR) == PSX_OK, PSX_ERR_ARGUMENT,
              "unknown covariance kind");
  PSX_REQUIRE(nblk >= 1, PSX_ERR_ARGUMENT, "at least one block");
  PSX_REQUIRE(dW && dC && dspectra && (nR == 0 || dratios), PSX_ERR_ARGUMENT, "null pointer");
  if (precision == PSX_FP64) {
    for (int b = 0; b < nblk; ++b)
      run_contraction(c, cov_request(kind, lmin, lmax, dspectra + b * nS,
                                     nR ? dratios + b * nR : nullptr, dW + b * nW, LW,
                                     precision, dC[b], ldC));
    return;
  }
  const int fam = mix_family_of_cov(kind);
  const int per = std::min(16, 16 * PSX_MIX_NCG_MAX / nW);  // blocks per matrix-core pass
  for (int b0 = 0; b0 < nblk; b0 += per) {
    const int nb = std::min(per, nblk - b0);
    MixRequest mr;
    mr.lmin = lmin;
    mr.lmax = lmax;
    mr.LW = LW;
    mr.fam = fam;
    mr.ncol = nb * nW;
    mr.epi = EPI_COV;
    mr.cov_kind = kind;
    mr.nblk = nb;
    mr.nW = nW;
    mr.ldo = ldC;
    mr.mode = precision;
    for (int b = 0; b < nb; ++b) {
      const int B = b0 + b;
      PSX_REQUIRE(dC[B], PSX_ERR_ARGUMENT, "null output");
      for (int w = 0; w < nW; ++w) {
        PSX_REQUIRE(dW[B * nW + w], PSX_ERR_ARGUMENT, "null W spectrum");
        mr.dW[b * nW + w] = dW[B * nW + w];
      }
      for (int i = 0; i < 4; ++i) {
        mr.dspec[b][i] = i < nS ? dspectra[B * nS + i] : nullptr;
        mr.dratio[b][i] = i < nR ? dratios[B * nR + i] : nullptr;
        if (i < nS) PSX_REQUIRE(mr.dspec[b][i], PSX_ERR_ARGUMENT, "null spectrum");
        if (i < nR) PSX_REQUIRE(mr.dratio[b][i], PSX_ERR_ARGUMENT, "null noise ratio");
      }
      mr.dout[b] = dC[B];
    }
    run_mixed(c, mr);
  }
}

}  // namespace

extern "C" {

int psx_cov_arity(int kind, int* nW, int* nspectra, int* nratios) {
  static const int tab[7][3] = {{8, 4, 4}, {8, 4, 4}, {4, 4, 2}, {5, 4, 2},
                                {4, 4, 2}, {4, 4, 2}, {2, 4, 0}};
  if (kind < 0 || kind > 6) return PSX_ERR_ARGUMENT;
  if (nW) *nW = tab[kind][0];
  if (nspectra) *nspectra = tab[kind][1];
  if (nratios) *nratios = tab[kind][2];
  return PSX_OK;
}

int psx_mcm_d(psx_ctx* h, int kind, int lmin, int lmax, const double* dW,
              int64_t LW, double* dM, int64_t ldM) {
  Ctx* c = C(h);
  return guarded(c, [&] {
    check_block(lmin, lmax, LW, ldM);
    PSX_REQUIRE(dW && dM, PSX_ERR_ARGUMENT, "null pointer");
    PSX_HIP(hipSetDevice(c->device));
    run_contraction(c, mcm_request(kind, lmin, lmax, dW, LW, dM, ldM));
  });
}

int psx_mcm_d_noalloc(psx_ctx* h, int kind, int lmin, int lmax, const double* dW,
                      int64_t LW, double* dM, int64_t ldM) {
  Ctx* c = C(h);
  return guarded(c, [&] {
    check_block(lmin, lmax, LW, ldM);
    PSX_REQUIRE(dW && dM, PSX_ERR_ARGUMENT, "null pointer");
    PSX_HIP(hipSetDevice(c->device));
    run_contraction_noalloc(c, mcm_request(kind, lmin, lmax, dW, LW, dM, ldM));
  });
}

int psx_mcm(psx_ctx* h, int kind, int lmin, int lmax, const double* W, int64_t LW,
            double* M, int64_t ldM) {
  Ctx* c = C(h);
  return guarded(c, [&] {
    check_block(lmin, lmax, LW, ldM);
    PSX_REQUIRE(W && M, PSX_ERR_ARGUMENT, "null pointer");
    PSX_HIP(hipSetDevice(c->device));
    const int n = lmax - lmin + 1;
    const size_t wb = sizeof(double) * LW, mb = sizeof(double) * n * (size_t)n;
    XferBuf x(c, xfer_bytes({wb, mb}));
    double* dW = x.take(wb);
    double* dM = x.take(mb);
    h2d(c, dW, W, wb);
    run_contraction_to_host(c, mcm_request(kind, lmin, lmax, dW, LW, dM, n), &M, ldM);
  });
}

int psx_mcm_binned(psx_ctx* h, int kind, int lmin, int lmax, const double* W, int64_t LW,
                   const double* P, int nbl, const double* Q, int nbr, double* out,
                   int64_t ldo) {
  Ctx* c = C(h);
  return guarded(c, [&] {
    check_block(lmin, lmax, LW, lmax - lmin + 1);
    PSX_REQUIRE(W && P && Q && out, PSX_ERR_ARGUMENT, "null pointer");
    PSX_REQUIRE(nbl >= 1 && nbr >= 1, PSX_ERR_ARGUMENT, "need at least one bin on each side");
    PSX_REQUIRE(ldo >= nbl, PSX_ERR_AXES, "leading dimension of the binned matrix below nbl");
    PSX_HIP(hipSetDevice(c->device));
    const int n = lmax - lmin + 1;
    const size_t wb = sizeof(double) * LW, mb = sizeof(double) * n * (size_t)n,
                 pb = sizeof(double) * nbl * (size_t)n, qb = sizeof(double) * n * (size_t)nbr,
                 ob = sizeof(double) * nbl * (size_t)nbr;
    XferBuf x(c, xfer_bytes({wb, mb, pb, qb, pb, ob}));
    double* dW = x.take(wb);
    double* dM = x.take(mb);
    double* dP = x.take(pb);
    double* dQ = x.take(qb);
    double* dT = x.take(pb);
    double* dO = x.take(ob);
    h2d(c, dW, W, wb);
    h2d_matrix(c, dP, nbl, P, nbl, nbl, n);
    h2d_matrix(c, dQ, n, Q, n, n, nbr);
    run_contraction(c, mcm_request(kind, lmin, lmax, dW, LW, dM, n));
    dgemm_dev(c, nbl, n, n, dP, nbl, dM, n, dT, nbl);     // P M
    dgemm_dev(c, nbl, nbr, n, dT, nbl, dQ, n, dO, nbl);   // (P M) Q
    d2h_matrix(c, out, ldo, dO, nbl, nbl, nbr);
  });
}

int psx_mcm_pm_d(psx_ctx* h, int lmin, int lmax, const double* dW, int64_t LW,
                 double* dMpp, double* dMmm, int64_t ldM) {
  Ctx* c = C(h);
  return guarded(c, [&] {
    check_block(lmin, lmax, LW, ldM);
    PSX_REQUIRE(dW && dMpp && dMmm, PSX_ERR_ARGUMENT, "null pointer");
    PSX_HIP(hipSetDevice(c->device));
    run_contraction(c, pm_request(lmin, lmax, dW, LW, dMpp, dMmm, ldM));
  });
}

int psx_mcm_pm(psx_ctx* h, int lmin, int lmax, const double* W, int64_t LW,
               double* Mpp, double* Mmm, int64_t ldM) {
  Ctx* c = C(h);
  return guarded(c, [&] {
    check_block(lmin, lmax, LW, ldM);
    PSX_REQUIRE(W && Mpp && Mmm, PSX_ERR_ARGUMENT, "null pointer");
    PSX_HIP(hipSetDevice(c->device));
    const int n = lmax - lmin + 1;
    const size_t wb = sizeof(double) * LW, mb = sizeof(double) * n * (size_t)n;
    XferBuf x(c, xfer_bytes({wb, mb, mb}));
    double* dW = x.take(wb);
    double* dP = x.take(mb);
    double* dQ = x.take(mb);
    h2d(c, dW, W, wb);
    double* const host[2] = {Mpp, Mmm};
    run_contraction_to_host(c, pm_request(lmin, lmax, dW, LW, dP, dQ, n), host, ldM);
  });
}

int psx_mcm_set_d(psx_ctx* h, int lmin, int lmax, int64_t LW, const double* dW_TT,
                  const double* dW_TE, const double* dW_ET, const double* dW_PP,
                  double* dM_TT, double* dM_TE, double* dM_ET, double* dM_PP,
                  double* dM_MM, int64_t ldM) {
  Ctx* c = C(h);
  return guarded(c, [&] {
    check_block(lmin, lmax, LW, ldM);
    PSX_HIP(hipSetDevice(c->device));
    const bool all = dW_TT && dW_TE && dW_ET && dW_PP;
    if (all) {
      PSX_REQUIRE(dM_TT && dM_TE && dM_ET && dM_PP && dM_MM, PSX_ERR_ARGUMENT,
                  "null output pointer");
      const double* const dW[4] = {dW_TT, dW_TE, dW_ET, dW_PP};
      double* const dM[5] = {dM_TT, dM_TE, dM_ET, dM_PP, dM_MM};
      run_contraction(c, set_request(lmin, lmax, LW, dW, dM, ldM));
      return;
    }
    // partial set: run the blocks that were asked for
    int st = PSX_OK;
    if (dW_TT && st == PSX_OK) st = psx_mcm_d(h, PSX_MCM_00, lmin, lmax, dW_TT, LW, dM_TT, ldM);
    if (dW_TE && st == PSX_OK) st = psx_mcm_d(h, PSX_MCM_02, lmin, lmax, dW_TE, LW, dM_TE, ldM);
    if (dW_ET && st == PSX_OK) st = psx_mcm_d(h, PSX_MCM_02, lmin, lmax, dW_ET, LW, dM_ET, ldM);
    if (dW_PP && st == PSX_OK)
      st = psx_mcm_pm_d(h, lmin, lmax, dW_PP, LW, dM_PP, dM_MM, ldM);
    if (st != PSX_OK) throw Error{st, c->last_error};
  });
}

int psx_mcm_set(psx_ctx* h, int lmin, int lmax, int64_t LW, const double* W_TT,
                const double* W_TE, const double* W_ET, const double* W_PP,
                double* M_TT, double* M_TE, double* M_ET, double* M_PP, double* M_MM,
                int64_t ldM) {
  Ctx* c = C(h);
  return guarded(c, [&] {
    check_block(lmin, lmax, LW, ldM);
    PSX_HIP(hipSetDevice(c->device));
    const int n = lmax - lmin + 1;
    const size_t wb = sizeof(double) * LW, mb = sizeof(double) * n * (size_t)n;
    const double* Wh[4] = {W_TT, W_TE, W_ET, W_PP};
    double* Mh[5] = {M_TT, M_TE, M_ET, M_PP, M_MM};
    const bool want[5] = {W_TT != nullptr, W_TE != nullptr, W_ET != nullptr,
                          W_PP != nullptr, W_PP != nullptr};
    for (int k = 0; k < 5; ++k)
      if (want[k]) PSX_REQUIRE(Mh[k] != nullptr, PSX_ERR_ARGUMENT, "null output pointer");
    XferBuf x(c, xfer_bytes({wb, wb, wb, wb, mb, mb, mb, mb, mb}));
    double* dWp[4];
    double* dMp[5];
    for (int k = 0; k < 4; ++k) {
      dWp[k] = x.take(wb);
      if (Wh[k]) h2d(c, dWp[k], Wh[k], wb);
    }
    for (int k = 0; k < 5; ++k) dMp[k] = x.take(mb);
    if (W_TT && W_TE && W_ET && W_PP) {
      run_contraction_to_host(c, set_request(lmin, lmax, LW, dWp, dMp, n), Mh, ldM);
      return;
    }
    // partial set: the requested blocks one after another
    for (int k = 0; k < 3; ++k)
      if (Wh[k]) {
        const int kind = k == 0 ? PSX_MCM_00 : PSX_MCM_02;
        run_contraction_to_host(c, mcm_request(kind, lmin, lmax, dWp[k], LW, dMp[k], n),
                                &Mh[k], ldM);
      }
    if (W_PP)
      run_contraction_to_host(c, pm_request(lmin, lmax, dWp[3], LW, dMp[3], dMp[4], n),
                              &Mh[3], ldM);
  });
}

int psx_master_decouple(psx_ctx* h, int lmin, int lmax, int64_t LW, const double* W_TT,
                        const double* W_TE, const double* W_ET, const double* W_PP,
                        const double* pcl, double* out) {
  Ctx* c = C(h);
  return guarded(c, [&] {
    check_block(lmin, lmax, LW, lmax - lmin + 1);
    PSX_REQUIRE(W_TT && W_TE && W_ET && W_PP && pcl && out, PSX_ERR_ARGUMENT, "null pointer");
    PSX_HIP(hipSetDevice(c->device));
    const int n = lmax - lmin + 1, N = 2 * n;
    PSX_REQUIRE((int64_t)N * N <= (int64_t(1) << 31) && N <= INT32_MAX / 2, PSX_ERR_ARGUMENT,
                "lmax - lmin too large for the EE/BB system");
    // Every block has leading dimension 2n: TT, TE and ET each in an 2n x n
    // buffer (top half used), ++ and -- straight into the left half of the
    // 2n x 2n EE/BB matrix [M++ M--; M-- M++].
    const size_t wb = sizeof(double) * LW, tb = sizeof(double) * (size_t)N * n,
                 ab = sizeof(double) * (size_t)N * N, bb = sizeof(double) * 9 * (size_t)n,
                 pb = sizeof(int) * ((size_t)N + 64);
    XferBuf x(c, xfer_bytes({wb, wb, wb, wb, tb, tb, tb, ab, bb, pb, pb, pb, pb}));
    const double* Wh[4] = {W_TT, W_TE, W_ET, W_PP};
    double* dW[4];
    for (int k = 0; k < 4; ++k) {
      dW[k] = x.take(wb);
      h2d(c, dW[k], Wh[k], wb);
    }
    double* dTT = x.take(tb);
    double* dTE = x.take(tb);
    double* dET = x.take(tb);
    double* dA = x.take(ab);
    double* dB = x.take(bb);
    int* piv[4];
    for (int k = 0; k < 4; ++k) piv[k] = reinterpret_cast<int*>(x.take(pb));
    double* const dM[5] = {dTT, dTE, dET, dA, dA + n};
    run_contraction(c, set_request(lmin, lmax, LW, dW, dM, N));
    const int64_t off = (int64_t)n * N + n;   // the bottom-right block
    PSX_HIP(hipMemcpy2DAsync(dA + off, sizeof(double) * N, dA, sizeof(double) * N,
                             sizeof(double) * n, n, hipMemcpyDeviceToDevice, c->stream));
    PSX_HIP(hipMemcpy2DAsync(dA + (int64_t)n * N, sizeof(double) * N, dA + n,
                             sizeof(double) * N, sizeof(double) * n, n,
                             hipMemcpyDeviceToDevice, c->stream));
    // right-hand sides, in the order of pcl (TT, TE, TB, ET, BT, EE, BB, EB,
    // BE) with BE negated: [EB; -BE] solves [M++ M--; M-- M++] when [EB; BE]
    // solves [M++ -M--; -M-- M++] = D [M++ M--; M-- M++] D, D = diag(I, -I)
    std::vector<double> rhs(pcl, pcl + 9 * (size_t)n);
    for (int i = 0; i < n; ++i) rhs[8 * (size_t)n + i] = -rhs[8 * (size_t)n + i];
    h2d(c, dB, rhs.data(), bb);
    const LuSystem sys[4] = {
        {N, dA, N, dB + 5 * (size_t)n, N, 2, piv[0]},   // EE, BB | EB, -BE
        {n, dTT, N, dB, n, 1, piv[1]},                  // TT
        {n, dTE, N, dB + n, n, 2, piv[2]},              // TE | TB
        {n, dET, N, dB + 3 * (size_t)n, n, 2, piv[3]},  // ET | BT
    };
    lu_solve_concurrent(c, sys, 4);
    PSX_HIP(hipMemcpyAsync(rhs.data(), dB, bb, hipMemcpyDeviceToHost, c->stream));
    PSX_HIP(hipStreamSynchronize(c->stream));
    for (int i = 0; i < n; ++i) rhs[8 * (size_t)n + i] = -rhs[8 * (size_t)n + i];
    std::copy(rhs.begin(), rhs.end(), out);
  });
}

int psx_cov_d(psx_ctx* h, int kind, int lmin, int lmax, const double* const* dspectra,
              const double* const* dratios, const double* const* dW, int64_t LW,
              int precision, double* dC, int64_t ldC) {
  Ctx* c = C(h);
  return guarded(c, [&] {
    check_block(lmin, lmax, LW, ldC);
    ContractRequest rq =
        cov_request(kind, lmin, lmax, dspectra, dratios, dW, LW, precision, dC, ldC);
    PSX_HIP(hipSetDevice(c->device));
    if (precision != PSX_FP64) {
      double* const outs[1] = {dC};
      cov_batch_dev(c, kind, lmin, lmax, 1, dspectra, dratios, dW, LW, precision, outs, ldC);
      return;
    }
    run_contraction(c, rq);
  });
}

int psx_cov(psx_ctx* h, int kind, int lmin, int lmax, const double* const* spectra,
            const double* const* ratios, const double* const* W, int64_t LW,
            int precision, double* Cm, int64_t ldC) {
  Ctx* c = C(h);
  return guarded(c, [&] {
    check_block(lmin, lmax, LW, ldC);
    int nW = 0, nS = 0, nR = 0;
    PSX_REQUIRE(psx_cov_arity(kind, &nW, &nS, &nR) == PSX_OK, PSX_ERR_ARGUMENT,
                "unknown covariance kind");
    PSX_REQUIRE(W && Cm && spectra, PSX_ERR_ARGUMENT, "null pointer");
    for (int i = 0; i < nS; ++i) PSX_REQUIRE(spectra[i], PSX_ERR_ARGUMENT, "null spectrum");
    for (int i = 0; i < nR; ++i)
      PSX_REQUIRE(ratios && ratios[i], PSX_ERR_ARGUMENT, "null noise ratio");
    for (int i = 0; i < nW; ++i) PSX_REQUIRE(W[i], PSX_ERR_ARGUMENT, "null W spectrum");
    PSX_HIP(hipSetDevice(c->device));
    const int n = lmax - lmin + 1;
    const size_t sb = sizeof(double) * (lmax + 1), wb = sizeof(double) * LW;
    const size_t cb = sizeof(double) * n * (size_t)n;
    XferBuf x(c, xfer_bytes({(nS + nR) * sb + nW * wb, cb}));
    char* p = (char*)x.take((nS + nR) * sb + nW * wb);
    double* dC = x.take(cb);
    const double* ds[4] = {nullptr, nullptr, nullptr, nullptr};
    const double* dr[4] = {nullptr, nullptr, nullptr, nullptr};
    const double* dw[8] = {nullptr};
    for (int i = 0; i < nS; ++i, p += sb) {
      h2d(c, p, spectra[i], sb);
      ds[i] = (const double*)p;
    }
    for (int i = 0; i < nR; ++i, p += sb) {
      h2d(c, p, ratios[i], sb);
      dr[i] = (const double*)p;
    }
    for (int i = 0; i < nW; ++i, p += wb) {
      h2d(c, p, W[i], wb);
      dw[i] = (const double*)p;
    }
    ContractRequest rq = cov_request(kind, lmin, lmax, ds, dr, dw, LW, precision, dC, n);
    if (precision != PSX_FP64) {
      double* const outs[1] = {dC};
      cov_batch_dev(c, kind, lmin, lmax, 1, ds, dr, dw, LW, precision, outs, n);
      d2h_matrix(c, Cm, ldC, dC, n, n, n);
      return;
    }
    run_contraction_to_host(c, rq, &Cm, ldC);
  });
}


int psx_mcm_batch_d(psx_ctx* h, int kind, int lmin, int lmax, int K, const double* const* dW,
                    int64_t LW, int precision, double* const* dM, int64_t ldM) {
  Ctx* c = C(h);
  return guarded(c, [&] {
    check_block(lmin, lmax, LW, ldM);
    check_precision(precision);
    PSX_HIP(hipSetDevice(c->device));
    mcm_batch_dev(c, kind, lmin, lmax, K, dW, LW, precision, dM, ldM);
  });
}

int psx_mcm_batch(psx_ctx* h, int kind, int lmin, int lmax, int K, const double* const* W,
                  int64_t LW, int precision, double* const* M, int64_t ldM) {
  Ctx* c = C(h);
  return guarded(c, [&] {
    check_block(lmin, lmax, LW, ldM);
    check_precision(precision);
    PSX_REQUIRE(K >= 1 && K <= 64, PSX_ERR_ARGUMENT, "between 1 and 64 W spectra");
    PSX_REQUIRE(W && M, PSX_ERR_ARGUMENT, "null pointer");
    for (int k = 0; k < K; ++k) PSX_REQUIRE(W[k] && M[k], PSX_ERR_ARGUMENT, "null pointer");
    PSX_HIP(hipSetDevice(c->device));
    const int n = lmax - lmin + 1;
    const size_t wb = sizeof(double) * LW, mb = sizeof(double) * n * (size_t)n;
    std::vector<size_t> pieces((size_t)2 * K);
    for (int k = 0; k < K; ++k) {
      pieces[k] = wb;
      pieces[K + k] = mb;
    }
    size_t total = 0;
    for (size_t b : pieces) total += (b + 255) & ~size_t(255);
    XferBuf x(c, total);
    std::vector<const double*> dW(K);
    std::vector<double*> dM(K);
    for (int k = 0; k < K; ++k) {
      double* d = x.take(wb);
      h2d(c, d, W[k], wb);
      dW[k] = d;
    }
    for (int k = 0; k < K; ++k) dM[k] = x.take(mb);
    mcm_batch_dev(c, kind, lmin, lmax, K, dW.data(), LW, precision, dM.data(), n);
    std::vector<D2HMatrix> out;
    for (int k = 0; k < K; ++k) out.push_back({M[k], ldM, dM[k], n, n, n});
    d2h_matrices(c, out.data(), K);
  });
}

int psx_cov_batch_d(psx_ctx* h, int kind, int lmin, int lmax, int nblk,
                    const double* const* dspectra, const double* const* dratios,
                    const double* const* dW, int64_t LW, int precision, double* const* dC,
                    int64_t ldC) {
  Ctx* c = C(h);
  return guarded(c, [&] {
    check_block(lmin, lmax, LW, ldC);
    check_precision(precision);
    PSX_HIP(hipSetDevice(c->device));
    cov_batch_dev(c, kind, lmin, lmax, nblk, dspectra, dratios, dW, LW, precision, dC, ldC);
  });
}

int psx_cov_batch(psx_ctx* h, int kind, int lmin, int lmax, int nblk,
                  const double* const* spectra, const double* const* ratios,
                  const double* const* W, int64_t LW, int precision, double* const* Cm,
                  int64_t ldC) {
  Ctx* c = C(h);
  return guarded(c, [&] {
    check_block(lmin, lmax, LW, ldC);
    check_precision(precision);
    int nW = 0, nS = 0, nR = 0;
    PSX_REQUIRE(psx_cov_arity(kind, &nW, &nS, &nR) == PSX_OK, PSX_ERR_ARGUMENT,
                "unknown covariance kind");
    PSX_REQUIRE(nblk >= 1 && nblk <= 64, PSX_ERR_ARGUMENT, "between 1 and 64 blocks");
    PSX_REQUIRE(spectra && W && Cm && (nR == 0 || ratios), PSX_ERR_ARGUMENT, "null pointer");
    for (int i = 0; i < nblk * nS; ++i) PSX_REQUIRE(spectra[i], PSX_ERR_ARGUMENT, "null spectrum");
    for (int i = 0; i < nblk * nR; ++i)
      PSX_REQUIRE(ratios[i], PSX_ERR_ARGUMENT, "null noise ratio");
    for (int i = 0; i < nblk * nW; ++i) PSX_REQUIRE(W[i], PSX_ERR_ARGUMENT, "null W spectrum");
    for (int b = 0; b < nblk; ++b) PSX_REQUIRE(Cm[b], PSX_ERR_ARGUMENT, "null output");
    PSX_HIP(hipSetDevice(c->device));
    const int n = lmax - lmin + 1;
    const size_t sb = sizeof(double) * (lmax + 1), wb = sizeof(double) * LW;
    const size_t cb = sizeof(double) * n * (size_t)n;
    const size_t total = (size_t)nblk * (nS + nR) * ((sb + 255) & ~size_t(255)) +
                         (size_t)nblk * nW * ((wb + 255) & ~size_t(255)) +
                         (size_t)nblk * ((cb + 255) & ~size_t(255));
    XferBuf x(c, total);
    std::vector<const double*> ds((size_t)nblk * nS), dr((size_t)nblk * nR),
        dw((size_t)nblk * nW);
    std::vector<double*> dc(nblk);
    for (size_t i = 0; i < ds.size(); ++i) {
      double* d = x.take(sb);
      h2d(c, d, spectra[i], sb);
      ds[i] = d;
    }
    for (size_t i = 0; i < dr.size(); ++i) {
      double* d = x.take(sb);
      h2d(c, d, ratios[i], sb);
      dr[i] = d;
    }
    for (size_t i = 0; i < dw.size(); ++i) {
      double* d = x.take(wb);
      h2d(c, d, W[i], wb);
      dw[i] = d;
    }
    for (int b = 0; b < nblk; ++b) dc[b] = x.take(cb);
    cov_batch_dev(c, kind, lmin, lmax, nblk, ds.data(), nR ? dr.data() : nullptr, dw.data(), LW,
                  precision, dc.data(), n);
    PSX_HIP(hipStreamSynchronize(c->stream));
    std::vector<D2HMatrix> out;
    for (int b = 0; b < nblk; ++b) out.push_back({Cm[b], ldC, dc[b], n, n, n});
    d2h_matrices(c, out.data(), nblk);
  });
}

}  // extern "C"
