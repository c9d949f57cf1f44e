"""CovField and CovarianceWorkspace: inputs and W-spectra cache of coupledcov.

Mirrors src/workspace.jl: CovField (:14-61), CovarianceWorkspace (:77-135),
effective_weight_alm! (:141-171) and window_function_W! (:174-213).  W spectra
are cross-spectra of effective weight maps, cached by the same 8-tuple key
(X, Y, i, j, alpha, p, q, beta) as the reference (WIndex, :74).
"""
from __future__ import annotations

from typing import Dict, Tuple

import numpy as np

from .alm import Alm, alm2cl, num_alms
from .sht import HealpixMap, map2alm, nside2lmax, pixsize
from .spectralarray import SpectralVector

SpectrumName = Tuple[str, str, str]   # (XY, field_i, field_j), reference src/workspace.jl:71
WIndex = Tuple[str, str, str, str, str, str, str, str]


class ConstantDict(dict):
    """A dictionary that returns one value for any key (reference src/util.jl:51-54)."""

    def __init__(self, value):
        super().__init__()
        self.c = value

    def __getitem__(self, key):
        return self.c

    def __len__(self):
        return 1


def _as_map(m) -> HealpixMap:
    return m if isinstance(m, HealpixMap) else HealpixMap(m)


class CovField:
    """A field: name, T and P masks, pixel variances (II, QQ, UU), beams."""

    def __init__(self, name: str, maskT, maskP, sigma2=None, beamT=None, beamP=None):
        self.name = name
        self.maskT = _as_map(maskT)
        self.maskP = _as_map(maskP)
        ns = self.maskT.nside
        if sigma2 is None:
            z = np.zeros(self.maskT.npix)
            sigma2 = (z, z, z)
        self.sigma2 = tuple(_as_map(s) for s in sigma2)
        self.beamT = SpectralVector(np.ones(3 * ns)) if beamT is None else beamT
        self.beamP = SpectralVector(np.ones(3 * ns)) if beamP is None else beamP


def _split_maptype(XY: str):
    """:TP -> (:TT, :PP) (src/workspace.jl:65-68)."""
    a, b = XY[0], XY[1]
    return a + a, b + b


class CovarianceWorkspace:
    """Inputs and cache for covariance calculations (src/workspace.jl:77-135).

    `map2alm` may be replaced (e.g. by a faster transform) through the
    `transform` argument: a callable (map, lmax) -> Alm."""

    def __init__(self, m_i: CovField, m_j: CovField, m_p: CovField, m_q: CovField,
                 lmax: int = 0, transform=None):
        self.field_names = (m_i.name, m_j.name, m_p.name, m_q.name)
        self.lmax = nside2lmax(m_i.maskT.nside) if lmax == 0 else int(lmax)
        self.mask_p = {}
        self.weight_p = {}
        for f in (m_i, m_j, m_p, m_q):
            self.mask_p[(f.name, "TT")] = f.maskT
            self.mask_p[(f.name, "PP")] = f.maskP
            self.weight_p[(f.name, "II")] = f.sigma2[0]
            self.weight_p[(f.name, "QQ")] = f.sigma2[1]
            self.weight_p[(f.name, "UU")] = f.sigma2[2]
        self.effective_weights: Dict[tuple, Alm] = {}
        self.W_spectra: Dict[WIndex, np.ndarray] = {}
        self._transform = transform or (lambda m, lmax: map2alm(m, lmax=lmax))

    def effective_weight_alm(self, A: str, i: str, j: str, alpha: str) -> Alm:
        """src/workspace.jl:141-171: SHT of m_i^X m_j^Y (A = ∅∅) or, when
        i == j, of m_i^X m_j^Y sigma^2_A Omega_p (A in II, QQ, UU); else zero."""
        key = (A, i, j, alpha)
        if key in self.effective_weights:
            return self.effective_weights[key]
        X, Y = _split_maptype(alpha)
        m_iX = self.mask_p[(i, X)]
        m_jY = self.mask_p[(j, Y)]
        if A == "∅∅":
            buf = HealpixMap(m_iX.pixels * m_jY.pixels, m_iX.nside)
            w = self._transform(buf, self.lmax)
            self.effective_weights[key] = w
            return w
        if A in ("II", "QQ", "UU") and i == j:
            buf = m_iX.pixels * m_jY.pixels
            buf = buf * self.weight_p[(i, A)].pixels * pixsize(m_iX)
            w = self._transform(HealpixMap(buf, m_iX.nside), self.lmax)
            self.effective_weights[key] = w
            return w
        return Alm(self.lmax, self.lmax, np.zeros(num_alms(self.lmax), dtype=np.complex128))

    def window_function_W(self, X, Y, i, j, alpha, p, q, beta) -> np.ndarray:
        """src/workspace.jl:174-213: mean over the weight terms of X and Y of
        alm2cl(w_ij^alpha, w_pq^beta)[0:lmax]; TT -> II, PP -> (QQ, UU)."""
        key = (X, Y, i, j, alpha, p, q, beta)
        if key in self.W_spectra:
            return self.W_spectra[key]
        terms = {"TT": ("II",), "PP": ("QQ", "UU")}
        tX = terms.get(X, (X,))
        tY = terms.get(Y, (Y,))
        out = np.zeros(self.lmax + 1)
        for wX in tX:
            for wY in tY:
                out += alm2cl(self.effective_weight_alm(wX, i, j, alpha),
                              self.effective_weight_alm(wY, p, q, beta))[: self.lmax + 1]
        out *= 1.0 / (len(tX) * len(tY))
        self.W_spectra[key] = out
        return out


SpectralWorkspace = CovarianceWorkspace  # name used by the reference's scripts
