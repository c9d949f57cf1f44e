"""mcm(): mode-coupling matrices on the GPU.

Mirrors the reference entry points (src/modecoupling.jl):
  mcm(spec, alm1, alm2; lmin, lmax)          :192-226
  mcm((:EE_BB, :EB_BE), alm1, alm2; ...)      :229-246
  mcm(spec, mask1::HealpixMap, mask2; ...)    :250-251 (map2alm of the masks)
  mcm(spec, f1::CovField, f2::CovField; ...)  :252-256 (T or P mask per letter)
The inner loops inner_mcm⁰⁰!/⁰²!/⁺⁺!/⁻⁻! (:78-159) run in libpsx_hip.so.
"""
from __future__ import annotations

import numpy as np

from ._abi import MCM_00, MCM_02, MCM_MM, MCM_PP, ArgumentError, AxisError, Context, default_context
from .alm import Alm, alm2cl
from .blockspectralmatrix import BlockSpectralMatrix, hvcat
from .spectralarray import SpectralArray, SpectralVector

# spectrum names accepted by the reference (src/modecoupling.jl:198-225);
# unicode and ASCII spellings of the block symbols
_KIND = {
    "TT": MCM_00, "M⁰⁰": MCM_00, "M00": MCM_00,
    "TE": MCM_02, "ET": MCM_02, "TB": MCM_02, "BT": MCM_02,
    "M⁰²": MCM_02, "M²⁰": MCM_02, "M02": MCM_02, "M20": MCM_02,
    "M⁺⁺": MCM_PP, "M++": MCM_PP, "Mpp": MCM_PP,
    "M⁻⁻": MCM_MM, "M--": MCM_MM, "Mmm": MCM_MM,
}
_BLOCKS = {"EE_BB", "EB_BE"}


def _name(spec) -> str:
    s = str(spec)
    return s[1:] if s.startswith(":") else s


def mcm_from_spectrum(spec, V, *, lmin: int = 0, lmax: int | None = None,
                      ctx: Context | None = None):
    """Mode-coupling matrix from the mask cross-spectrum V (0-based).

    The reference passes V = alm2cl(alm1, alm2)[0:lmax] (src/modecoupling.jl:197):
    W entries above lmax are dropped before the l3 sum."""
    ctx = ctx or default_context()
    V = np.asarray(V.parent if isinstance(V, SpectralArray) else V, dtype=np.float64)
    if lmax is None:
        lmax = V.size - 1
    if V.size < lmax + 1:
        raise ArgumentError(f"mask spectrum has {V.size} entries, lmax {lmax} needs {lmax + 1}")
    V = np.ascontiguousarray(V[: lmax + 1])
    name = _name(spec) if not isinstance(spec, tuple) else None
    firsts = [lmin, lmin]
    if isinstance(spec, tuple):
        if tuple(_name(s) for s in spec) != ("EE_BB", "EB_BE"):
            return None  # the reference returns nothing for other tuples (:235-246)
        P, M = ctx.mcm_pm(lmin, lmax, V)
        Mpp = SpectralArray(P, firsts=firsts)
        Mmm = SpectralArray(M, firsts=firsts)
        return (hvcat([[Mpp, Mmm], [Mmm, Mpp]]), hvcat([[Mpp, -Mmm], [-Mmm, Mpp]]))
    if name in _KIND:
        return SpectralArray(ctx.mcm(_KIND[name], lmin, lmax, V), firsts=firsts)
    if name in _BLOCKS:
        P, M = ctx.mcm_pm(lmin, lmax, V)
        Mpp = SpectralArray(P, firsts=firsts)
        Mmm = SpectralArray(M, firsts=firsts)
        if name == "EE_BB":
            return hvcat([[Mpp, Mmm], [Mmm, Mpp]])
        return hvcat([[Mpp, -Mmm], [-Mmm, Mpp]])
    raise ArgumentError(f"{spec} not a valid spectrum.")


def mcm_binned(spec, V, P, Q=None, *, lmin: int = 0, lmax: int | None = None,
               ctx: Context | None = None) -> np.ndarray:
    """P M Q for the single block M = mcm_from_spectrum(spec, V) (TT, TE/ET/TB/BT,
    M++ or M--), with M kept on the device.  P is a binning matrix as
    binning_matrix (src/util.jl:3-18) returns it, one column per l = 0..lmax;
    Q (default P^T) has one row per l.  Columns of P and rows of Q below lmin
    are dropped, as M starts at lmin."""
    ctx = ctx or default_context()
    V = np.asarray(V.parent if isinstance(V, SpectralArray) else V, dtype=np.float64)
    if lmax is None:
        lmax = V.size - 1
    if V.size < lmax + 1:
        raise ArgumentError(f"mask spectrum has {V.size} entries, lmax {lmax} needs {lmax + 1}")
    name = _name(spec)
    if name not in _KIND:
        raise ArgumentError(f"{spec}: binned matrices are for single blocks")
    P = np.asarray(P, dtype=np.float64)
    Q = P.T if Q is None else np.asarray(Q, dtype=np.float64)
    if P.ndim != 2 or Q.ndim != 2 or P.shape[1] != lmax + 1 or Q.shape[0] != lmax + 1:
        raise AxisError(f"binning matrices need lmax + 1 = {lmax + 1} columns (P) and rows (Q)")
    return ctx.mcm_binned(_KIND[name], lmin, lmax, np.ascontiguousarray(V[: lmax + 1]),
                          P[:, lmin:], Q[lmin:, :])


def mcm(spec, a1, a2=None, *, lmin: int = 0, lmax: int | None = None,
        ctx: Context | None = None):
    """mcm(spec, alm1, alm2), mcm(spec, mask1, mask2) or mcm(spec, field1, field2).

    Alm inputs: lmax defaults to min(alm lmax) and V = alm2cl(alm1, alm2)[0:lmax].
    HEALPix mask maps (ring-ordered numpy arrays or HealpixMap) are transformed
    with map2alm first; CovField inputs select maskT/maskP by the letters of
    spec."""
    from .workspace import CovField  # local import: workspace imports this module
    a2 = a1 if a2 is None else a2
    if isinstance(a1, CovField):
        name = _name(spec)
        m1 = a1.maskT if name[0] == "T" else a1.maskP
        m2 = a2.maskT if name[1] == "T" else a2.maskP
        return mcm(spec, m1, m2, lmin=lmin, lmax=lmax, ctx=ctx)
    if not isinstance(a1, Alm):
        # masks: map2alm (Healpix.jl defaults: lmax = 3 nside - 1, 3 iterations)
        # and their cross-spectrum in one device call, the alms never leave it
        from ._abi import default_context
        from .sht import HealpixMap, nside2lmax
        m1 = a1 if isinstance(a1, HealpixMap) else HealpixMap(a1)
        m2 = a2 if isinstance(a2, HealpixMap) else HealpixMap(a2)
        if m1.nside != m2.nside:
            raise ArgumentError("the two masks need the same nside")
        ctx = ctx or default_context()
        la = nside2lmax(m1.nside)
        C = ctx.mask_spectra(np.stack([m1.pixels, m2.pixels]), m1.nside, la, 3)
        lmax = la if lmax is None else lmax
        return mcm_from_spectrum(spec, C[0, 1][: lmax + 1], lmin=lmin, lmax=lmax, ctx=ctx)
    if lmax is None:
        lmax = min(a1.lmax, a2.lmax)
    V = alm2cl(a1, a2)[: lmax + 1]
    return mcm_from_spectrum(spec, V, lmin=lmin, lmax=lmax, ctx=ctx)


def mcm_set(W_TT, W_TE, W_ET, W_PP, *, lmin: int = 0, lmax: int | None = None,
            ctx: Context | None = None):
    """The five blocks maskedalm2spectra needs (TT, TE, ET, ++, --) in one
    fused pass (src/modecoupling.jl:348-377)."""
    ctx = ctx or default_context()
    Ws = [None if w is None else np.asarray(w, dtype=np.float64) for w in (W_TT, W_TE, W_ET, W_PP)]
    if lmax is None:
        lmax = min(w.size for w in Ws if w is not None) - 1
    Ws = [None if w is None else np.ascontiguousarray(w[: lmax + 1]) for w in Ws]
    outs = ctx.mcm_set(lmin, lmax, *Ws)
    return tuple(None if o is None else SpectralArray(o, firsts=[lmin, lmin]) for o in outs)


def scale(m, sT, sP=None):
    """scale!(m, s) / scale!(m, sT, sP) on a map or an (I, Q, U) triple (src/modecoupling.jl:259-268)."""
    if isinstance(m, (tuple, list)) and len(m) == 3:
        sP = sT if sP is None else sP
        m[0][...] *= sT
        m[1][...] *= sP
        m[2][...] *= sP
        return m
    m[...] *= sT
    return m


def mask(m, maskT, maskP=None):
    """mask!(m, mask) / mask!(polmap, maskT, maskP) (src/modecoupling.jl:283-296)."""
    if isinstance(m, (tuple, list)) and len(m) == 3:
        maskP = maskT if maskP is None else maskP
        m[0][...] *= np.asarray(maskT)
        m[1][...] *= np.asarray(maskP)
        m[2][...] *= np.asarray(maskP)
        return m
    m[...] *= np.asarray(maskT)
    return m
