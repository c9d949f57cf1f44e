"""coupledcov / decouple_covmat on the GPU.

Mirrors src/covariance.jl: `coupledcov(ch1, ch2, workspace, spectra,
noiseratios; lmin, lmax)` (:34-61) with the per-pair selection of spectra,
noise ratios and W spectra of coupledcovTTTT!/EEEE!/TTTE!/TETE!/TEEE!/TTEE!
(:64-418); the inner loops (:92-446) run in libpsx_hip.so (psx_cov).
"""
from __future__ import annotations

import numpy as np

from ._abi import (COV_EEEE, COV_TEEE, COV_TEEE_PLANCK, COV_TETE, COV_TTEE, COV_TTTE, COV_TTTT,
                   FP64, Context, default_context)
from .spectralarray import SpectralArray, parent, spectralones
from .workspace import ConstantDict, CovarianceWorkspace

E = "∅∅"


def _vec(x, lmax):
    a = np.asarray(parent(x), dtype=np.float64)
    return np.ascontiguousarray(a[: lmax + 1])


def _run(kind, C_lmin, lmax, spectra, ratios, W, ctx, precision):
    ctx = ctx or default_context()
    Ws = [np.ascontiguousarray(np.asarray(parent(w), dtype=np.float64)) for w in W]
    out = ctx.cov(kind, C_lmin, lmax, [_vec(s, lmax) for s in spectra],
                  [_vec(r, lmax) for r in ratios], Ws, precision)
    return SpectralArray(out, firsts=[C_lmin, C_lmin])


def coupledcovTTTT(ws: CovarianceWorkspace, spectra, noiseratios, lmin, lmax, ctx=None,
                   precision=FP64):
    """src/covariance.jl:64-88."""
    i, j, p, q = ws.field_names
    r = noiseratios
    W = ws.window_function_W
    return _run(COV_TTTT, lmin, lmax,
                [spectra["TT", i, p], spectra["TT", j, q], spectra["TT", i, q], spectra["TT", j, p]],
                [r["TT", i, p], r["TT", j, q], r["TT", i, q], r["TT", j, p]],
                [W(E, E, i, p, "TT", j, q, "TT"), W(E, E, i, q, "TT", j, p, "TT"),
                 W(E, "TT", i, p, "TT", j, q, "TT"), W(E, "TT", j, q, "TT", i, p, "TT"),
                 W(E, "TT", i, q, "TT", j, p, "TT"), W(E, "TT", j, p, "TT", i, q, "TT"),
                 W("TT", "TT", i, p, "TT", j, q, "TT"), W("TT", "TT", i, q, "TT", j, p, "TT")],
                ctx, precision)


def coupledcovEEEE(ws, spectra, noiseratios, lmin, lmax, ctx=None, precision=FP64):
    """src/covariance.jl:125-149."""
    i, j, p, q = ws.field_names
    r = noiseratios
    W = ws.window_function_W
    return _run(COV_EEEE, lmin, lmax,
                [spectra["EE", i, p], spectra["EE", j, q], spectra["EE", i, q], spectra["EE", j, p]],
                [r["EE", i, p], r["EE", j, q], r["EE", i, q], r["EE", j, p]],
                [W(E, E, i, p, "PP", j, q, "PP"), W(E, E, i, q, "PP", j, p, "PP"),
                 W(E, "PP", i, p, "PP", j, q, "PP"), W(E, "PP", j, q, "PP", i, p, "PP"),
                 W(E, "PP", i, q, "PP", j, p, "PP"), W(E, "PP", j, p, "PP", i, q, "PP"),
                 W("PP", "PP", i, p, "PP", j, q, "PP"), W("PP", "PP", i, q, "PP", j, p, "PP")],
                ctx, precision)


def coupledcovTTTE(ws, spectra, noiseratios, lmin, lmax, ctx=None, precision=FP64):
    """src/covariance.jl:186-204."""
    i, j, p, q = ws.field_names
    r = noiseratios
    W = ws.window_function_W
    return _run(COV_TTTE, lmin, lmax,
                [spectra["TT", i, p], spectra["TT", j, p], spectra["TE", i, q], spectra["TE", j, q]],
                [r["TT", i, p], r["TT", j, p]],
                [W(E, E, i, p, "TT", j, q, "TP"), W(E, E, i, q, "TP", j, p, "TT"),
                 W(E, "TT", j, q, "TP", i, p, "TT"), W(E, "TT", i, q, "TP", j, p, "TT")],
                ctx, precision)


def coupledcovTETE(ws, spectra, noiseratios, lmin, lmax, ctx=None, precision=FP64):
    """src/covariance.jl:238-257."""
    i, j, p, q = ws.field_names
    r = noiseratios
    W = ws.window_function_W
    return _run(COV_TETE, lmin, lmax,
                [spectra["TT", i, p], spectra["EE", j, q], spectra["TE", i, q], spectra["TE", j, p]],
                [r["TT", i, p], r["EE", j, q]],
                [W(E, E, i, p, "TT", j, q, "PP"), W(E, E, i, q, "TP", j, p, "PT"),
                 W(E, "PP", i, p, "TT", j, q, "PP"), W(E, "TT", j, q, "PP", i, p, "TT"),
                 W("TT", "PP", i, p, "TT", j, q, "PP")],
                ctx, precision)


def coupledcovTEEE(ws, spectra, noiseratios, lmin, lmax, ctx=None, precision=FP64,
                   planck: bool = True):
    """src/covariance.jl:305-333 (planck=True is the reference default)."""
    i, j, p, q = ws.field_names
    r = noiseratios
    W = ws.window_function_W
    return _run(COV_TEEE_PLANCK if planck else COV_TEEE, lmin, lmax,
                [spectra["EE", j, q], spectra["EE", j, p], spectra["TE", i, p], spectra["TE", i, q]],
                [r["EE", j, q], r["EE", j, p]],
                [W(E, E, i, p, "TP", j, q, "PP"), W(E, E, i, q, "TP", j, p, "PP"),
                 W(E, "PP", i, p, "TP", j, q, "PP"), W(E, "PP", i, q, "TP", j, p, "PP")],
                ctx, precision)


def coupledcovTTEE(ws, spectra, noiseratios, lmin, lmax, ctx=None, precision=FP64):
    """src/covariance.jl:405-418."""
    i, j, p, q = ws.field_names
    W = ws.window_function_W
    return _run(COV_TTEE, lmin, lmax,
                [spectra["TE", i, p], spectra["TE", i, q], spectra["TE", j, q], spectra["TE", j, p]],
                [],
                [W(E, E, i, p, "TP", j, q, "TP"), W(E, E, i, q, "TP", j, p, "TP")],
                ctx, precision)


def _name(ch) -> str:
    s = str(ch)
    return s[1:] if s.startswith(":") else s


def coupledcov(ch1, ch2, workspace: CovarianceWorkspace, spectra, noiseratios=None, *,
               lmin: int = 0, lmax: int | None = None, ctx: Context | None = None,
               precision: int = FP64):
    """Coupled covariance block between spectra ch1 and ch2 (src/covariance.jl:34-61).

    `spectra` and `noiseratios` are dicts keyed by (XY, field_i, field_j) with
    0-based spectra; missing noise ratios default to ones (ConstantDict).
    Unsupported pairs print a message and return None, as the reference does."""
    lmax = workspace.lmax if lmax is None else int(lmax)
    if noiseratios is None or len(noiseratios) == 0:
        noiseratios = ConstantDict(spectralones(range(0, lmax + 1)).parent)
    a, b = _name(ch1), _name(ch2)
    args = (workspace, spectra, noiseratios, lmin, lmax, ctx, precision)
    if (a, b) == ("TT", "TT"):
        return coupledcovTTTT(*args)
    if (a, b) == ("EE", "EE"):
        return coupledcovEEEE(*args)
    if (a, b) == ("TE", "TE"):
        return coupledcovTETE(*args)
    if (a, b) == ("TT", "TE"):
        return coupledcovTTTE(*args)
    if (a, b) == ("TT", "EE"):
        return coupledcovTTEE(*args)
    if (a, b) == ("TE", "EE"):
        return coupledcovTEEE(*args)
    print(f"{a},{b} not implemented")
    return None


def decouple_covmat(Y, B1, B2, ctx: Context | None = None) -> SpectralArray:
    """B1^{-1} Y B2^{-T} (src/covariance.jl:8-14); Y is not modified."""
    ctx = ctx or default_context()
    out = ctx.decouple_covmat(parent(Y), parent(B1), parent(B2))
    if isinstance(Y, SpectralArray):
        return SpectralArray(out, firsts=[Y.firstindex(0), Y.firstindex(1)])
    return SpectralArray(out)
