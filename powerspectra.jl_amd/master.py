"""Decoupled spectra from masked alms, and binning.

Mirrors src/modecoupling.jl:343-380 (maskedalm2spectra) and :317-341
(master), and src/util.jl:3-18 (binning_matrix).  The five mode-coupling
blocks come from one fused pass (TT, TE (= TB), ET (= BT), ++, --), the
solves from the GPU LU: four factorisations (TT; TE for TE and TB; ET for ET
and BT; the 2n x 2n EE/BB matrix for EE/BB and EB/BE) where the reference
factors seven matrices, all in one device call that returns only the spectra.
"""
from __future__ import annotations

import numpy as np

from ._abi import ArgumentError, Context, default_context
from .alm import Alm, alm2cl
from .spectralarray import SpectralArray


def _cross_spectra(sets, ctx: Context):
    """Function (i, j) -> alm2cl(sets[i], sets[j]) over a list of Alm.  Sets
    that are the same object are transformed once; sets of one lmax with
    mmax = lmax go through one device call (psx_alm2cl)."""
    uniq, slot = [], []
    for a in sets:
        k = next((u for u, b in enumerate(uniq) if b is a), None)
        if k is None:
            uniq.append(a)
            k = len(uniq) - 1
        slot.append(k)
    if len({(a.lmax, a.mmax) for a in uniq}) == 1 and uniq[0].mmax == uniq[0].lmax:
        cl = ctx.alm2cl([a.alm for a in uniq], uniq[0].lmax)
        return lambda i, j: cl[slot[i], slot[j]]
    return lambda i, j: alm2cl(sets[i], sets[j])


_SPECTRA = ("TT", "TE", "TB", "ET", "BT", "EE", "BB", "EB", "BE")


def _decoupled(W, P, lmin: int, lmax: int, ctx: Context) -> dict:
    """The nine spectra of maskedalm2spectra (src/modecoupling.jl:343-380) from
    the mask cross-spectra W(a, b) (a in "T1", "P1"; b in "T2", "P2") and the
    pseudo-spectra P(X, Y) of map 1's X with map 2's Y.

    One device call (psx_master_decouple) builds the five blocks of mcm_set
    and solves M \\ pCl for each spectrum (src/modecoupling.jl:355-377): TE and
    TB share mcm(TE) = mcm(TB), ET and BT share mcm(ET), one factorisation
    each with two right-hand sides; EE/BB and EB/BE are decoupled jointly by
    [M++ M--; M-- M++] and [M++ -M--; -M-- M++].  The second is D A D for the
    first, A, with D = diag(I, -I): [EB; BE] solves it exactly when [EB; -BE]
    solves A against [pEB; -pBE].  Sign flips of rows and columns change
    neither partial pivoting's choices nor any magnitude, so one factorisation
    of A serves both systems.  The four factorisations run at the same time
    on the device."""
    Ws = [np.asarray(W(a, b), dtype=np.float64)[: lmax + 1]
          for a, b in (("T1", "T2"), ("T1", "P2"), ("P1", "T2"), ("P1", "P2"))]
    # SpectralVector(alm2cl(a, b))[IdentityRange(lmin:end)] (src/modecoupling.jl:355)
    pcl = np.stack([np.asarray(P(XY[0], XY[1]), dtype=np.float64)[lmin: lmax + 1]
                    for XY in _SPECTRA])
    X = ctx.master_decouple(lmin, lmax, *Ws, pcl)
    return {XY: SpectralArray(np.ascontiguousarray(X[k]), firsts=(lmin,))
            for k, XY in enumerate(_SPECTRA)}


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
    masks = _cross_spectra([maskT1, maskP1, maskT2, maskP2], ctx)
    maps = _cross_spectra([T1, E1, B1, T2, E2, B2], ctx)
    wi = {"T1": 0, "P1": 1, "T2": 2, "P2": 3}
    pi = {"T": 0, "E": 1, "B": 2}
    return _decoupled(lambda a, b: masks(wi[a], wi[b]), lambda X, Y: maps(pi[X], 3 + pi[Y]),
                      lmin, lmax, ctx)


def master(map1, maskT1, maskP1, map2, maskT2, maskP2, *, already_masked: bool = False,
           lmin: int = 0, transform=None, ctx: Context | None = None) -> dict:
    """Decoupled spectra of two polarised maps (I, Q, U) with their masks
    (src/modecoupling.jl:317-341).  `transform` = (pol_map2alm, map2alm):
    the polarised and scalar analyses to use (default: the GPU transforms at
    lmax = 3 nside - 1 with three refinements: masking, both analyses and the
    spectra in one device call, psx_master_spectra, each distinct mask
    transformed once)."""
    from .sht import HealpixMap, nside2lmax
    px = lambda m: np.asarray(m.pixels if isinstance(m, HealpixMap) else m, dtype=np.float64)
    mT1, mP1, mT2, mP2 = maskT1, maskP1, maskT2, maskP2
    if transform is not None:
        pol, scal = transform
        I1, Q1, U1 = (px(m) for m in map1)
        I2, Q2, U2 = (px(m) for m in map2)
        if not already_masked:
            I1, Q1, U1 = I1 * px(mT1), Q1 * px(mP1), U1 * px(mP1)
            I2, Q2, U2 = I2 * px(mT2), Q2 * px(mP2), U2 * px(mP2)
        return maskedalm2spectra(pol(I1, Q1, U1), scal(px(mT1)), scal(px(mP1)),
                                 pol(I2, Q2, U2), scal(px(mT2)), scal(px(mP2)),
                                 lmin=lmin, ctx=ctx)
    ctx = ctx or default_context()
    nside = HealpixMap(px(map1[0])).nside
    pcl, W = ctx.master_spectra([px(x) for x in (*map1, *map2)],
                                [px(w) for w in (mT1, mP1, mT2, mP2)], nside, nside2lmax(nside),
                                3, already_masked)
    wi = {"T1": 0, "P1": 1, "T2": 0, "P2": 1}
    pi = {"T": 0, "E": 1, "B": 2}
    return _decoupled(lambda a, b: W[wi[a], wi[b]], lambda X, Y: pcl[pi[X], pi[Y]],
                      lmin, lmax=nside2lmax(nside), ctx=ctx)


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
