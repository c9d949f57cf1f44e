"""Decoupled spectra from masked alms, and binning.

Mirrors src/modecoupling.jl:343-380 (maskedalm2spectra) and :317-341
(master), and src/util.jl:3-18 (binning_matrix).  The five mode-coupling
blocks come from one fused pass (psx_mcm_set: TT, TE (= TB), ET (= BT), ++,
--), the solves from the GPU LU.
"""
from __future__ import annotations

import numpy as np

from ._abi import ArgumentError, Context, default_context
from .alm import Alm, alm2cl
from .blockspectralmatrix import block_vector, hvcat, solve, spectra
from .modecoupling import mcm_set
from .spectralarray import SpectralArray, SpectralVector


def _pcl(a: Alm, b: Alm, lmin: int, lmax: int) -> SpectralArray:
    """SpectralVector(alm2cl(a, b))[IdentityRange(lmin:end)] (src/modecoupling.jl:360)."""
    cl = alm2cl(a, b)[: lmax + 1]
    return SpectralVector(cl[lmin:], lmin - 1)


def maskedalm2spectra(maskedmap1, maskT1: Alm, maskP1: Alm, maskedmap2, maskT2: Alm,
                      maskP2: Alm, *, lmin: int = 0, ctx: Context | None = None) -> dict:
    """Spectra TT, TE, ET, TB, BT, EE, BB, EB, BE from the (T, E, B) alms of two
    masked maps and the alms of their masks (src/modecoupling.jl:343-380).

    Each independent spectrum is M \\ pCl with M = mcm(XY, maskX, maskY) (T mask
    for T, P mask for E and B); EE/BB and EB/BE are decoupled jointly with
    [M++ M--; M-- M++] and [M++ -M--; -M-- M++]."""
    ctx = ctx or default_context()
    T1, E1, B1 = maskedmap1
    T2, E2, B2 = maskedmap2
    lmax = min(a.lmax for a in (T1, E1, B1, T2, E2, B2, maskT1, maskP1, maskT2, maskP2))
    W = lambda a, b: alm2cl(a, b)[: lmax + 1]
    MTT, MTE, MET, Mpp, Mmm = mcm_set(W(maskT1, maskT2), W(maskT1, maskP2), W(maskP1, maskT2),
                                      W(maskP1, maskP2), lmin=lmin, lmax=lmax, ctx=ctx)
    first = dict(T=(T1, T2), E=(E1, E2), B=(B1, B2))
    blocks = {"TT": MTT, "TE": MTE, "ET": MET, "TB": MTE, "BT": MET}
    out = {}
    for XY, M in blocks.items():
        pcl = _pcl(first[XY[0]][0], first[XY[1]][1], lmin, lmax)
        out[XY] = solve(M, pcl, ctx=ctx)
    M_EE_BB = hvcat([[Mpp, Mmm], [Mmm, Mpp]])
    M_EB_BE = hvcat([[Mpp, -Mmm], [-Mmm, Mpp]])
    out["EE"], out["BB"] = spectra(solve(M_EE_BB, block_vector(_pcl(E1, E2, lmin, lmax),
                                                              _pcl(B1, B2, lmin, lmax)), ctx=ctx))
    out["EB"], out["BE"] = spectra(solve(M_EB_BE, block_vector(_pcl(E1, B2, lmin, lmax),
                                                              _pcl(B1, E2, lmin, lmax)), ctx=ctx))
    return out


def master(map1, maskT1, maskP1, map2, maskT2, maskP2, *, already_masked: bool = False,
           lmin: int = 0, transform=None, ctx: Context | None = None) -> dict:
    """Decoupled spectra of two polarised maps (I, Q, U) with their masks
    (src/modecoupling.jl:317-341).  `transform` = (pol_map2alm, map2alm):
    the polarised and scalar analyses to use (defaults: the GPU transforms,
    spin 0 and 2 for the maps, spin 0 for the masks)."""
    from .sht import HealpixMap, map2alm, map2alm_pol
    pol, scal = transform if transform is not None else (map2alm_pol, map2alm)
    px = lambda m: np.asarray(m.pixels if isinstance(m, HealpixMap) else m, dtype=np.float64)
    I1, Q1, U1 = (px(m) for m in map1)
    I2, Q2, U2 = (px(m) for m in map2)
    if not already_masked:
        I1, Q1, U1 = I1 * px(maskT1), Q1 * px(maskP1), U1 * px(maskP1)
        I2, Q2, U2 = I2 * px(maskT2), Q2 * px(maskP2), U2 * px(maskP2)
    return maskedalm2spectra(pol(I1, Q1, U1), scal(px(maskT1)), scal(px(maskP1)),
                             pol(I2, Q2, U2), scal(px(maskT2)), scal(px(maskP2)),
                             lmin=lmin, ctx=ctx)


def binning_matrix(left_bins, right_bins, weight_function, lmax: int | None = None) -> np.ndarray:
    """P[b, l] = w(l) / sum_{l' in bin b} w(l') for left_b <= l <= right_b,
    over the bins whose right edge is <= lmax (src/util.jl:3-18)."""
    left_bins = np.asarray(left_bins, dtype=np.int64)
    right_bins = np.asarray(right_bins, dtype=np.int64)
    if left_bins.shape != right_bins.shape:
        raise ArgumentError("left and right bin edges differ in length")
    lmax = int(right_bins[-1]) if lmax is None else int(lmax)
    # the reference cuts the left edges only (its `right_bings` typo keeps all
    # right edges); bins are then indexed 1..nbins on both arrays alike
    cut = right_bins <= lmax
    left_bins = left_bins[cut]
    nbins = left_bins.size
    P = np.zeros((nbins, lmax + 1))
    for b in range(nbins):
        ells = np.arange(left_bins[b], right_bins[b] + 1)
        w = np.array([weight_function(l) for l in ells], dtype=np.float64)
        P[b, left_bins[b]: right_bins[b] + 1] = w / w.sum()
    return P
