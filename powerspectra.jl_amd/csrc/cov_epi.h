// Covariance epilogues shared by the contraction kernels: the coefficient
// algebra of the reference's loop_cov*! bodies applied to one element's
// projector values xi[] (family order), with the spectra S[] and noise
// ratios R[] of the block in the reference loop's argument order.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/powerspectra_hip.h"

namespace psx {

// xi[] is in family order; see cov_layout() for the W -> xi mapping.
__device__ __forceinline__ double sqc(double x, bool& bad) {
  if (x < 0) bad = true;
  return sqrt(x);
}

template <int COVK>
__device__ __forceinline__ double cov_combine(int a, int b, const double* xi,
                                              const double* const* S, const double* const* R,
                                              bool& bad) {
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

}  // namespace psx
