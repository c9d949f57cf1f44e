"""Seeded cut-sky masks: a galactic cut |b| < b_cut plus circular holes at
uniform random positions, both cosine-apodised.  These are the masks the
mode-coupling and covariance workloads are specified on (cut 20 deg, holes of
radius 0.5 deg, 1 deg apodisation; nside 256 / 50 holes, nside 1024 / 200
holes), built ring by ring on the HEALPix grid of sht.RingInfo.
"""
from __future__ import annotations

import numpy as np

from .sht import RingInfo


def _taper(x: np.ndarray, width: float) -> np.ndarray:
    """0 for x <= 0, 1 for x >= width, 0.5 (1 - cos(pi x / width)) between."""
    t = np.clip(x / width, 0.0, 1.0)
    return 0.5 * (1.0 - np.cos(np.pi * t))


def cut_sky_mask(nside: int, nholes: int, seed: int, *, cut_deg: float = 20.0,
                 hole_deg: float = 0.5, apod_deg: float = 1.0) -> np.ndarray:
    """Ring-ordered mask: 0 for |b| < cut, rising to 1 over `apod_deg`; each
    hole is 0 within `hole_deg` of its centre and rises to 1 over `apod_deg`.
    Hole centres are uniform on the sphere, drawn from `seed`."""
    info = RingInfo(nside)
    npix = 12 * nside * nside
    cut, hole, apod = np.radians(cut_deg), np.radians(hole_deg), np.radians(apod_deg)
    theta_r = np.arccos(info.cth)
    mask = np.empty(npix)
    for k in range(info.first.size):
        b = np.pi / 2 - theta_r[k]
        mask[info.first[k]: info.first[k] + info.count[k]] = _taper(abs(b) - cut, apod)
    rng = np.random.default_rng(seed)
    z0 = rng.uniform(-1.0, 1.0, nholes)
    p0 = rng.uniform(0.0, 2 * np.pi, nholes)
    th0 = np.arccos(z0)
    reach = hole + apod
    for h in range(nholes):
        near = np.nonzero(np.abs(theta_r - th0[h]) <= reach)[0]
        for k in near:
            n = int(info.count[k])
            phi = info.phi0[k] + 2 * np.pi * np.arange(n) / n
            cosd = info.cth[k] * z0[h] + info.sth[k] * np.sqrt(1 - z0[h] ** 2) * np.cos(phi - p0[h])
            d = np.arccos(np.clip(cosd, -1.0, 1.0))
            sel = d < reach
            if sel.any():
                sl = slice(info.first[k], info.first[k] + n)
                seg = mask[sl]
                seg[sel] *= _taper(d[sel] - hole, apod)
    return mask


def config_masks(nside: int, nholes: int, seeds) -> list:
    """The workloads' masks for the given seeds (one mask per seed)."""
    return [cut_sky_mask(nside, nholes, s) for s in seeds]


def mask_spectra(masks, lmax: int, ctx=None, niter: int = 3) -> dict:
    """Cross-spectra W[i, j] = alm2cl(map2alm(m_i), map2alm(m_j))[0:lmax] of
    a list of masks (GPU transform and cross-spectra in one call per 8 masks,
    the alms stay on the device; nside taken from the masks)."""
    from ._abi import default_context
    ctx = ctx or default_context()
    M = np.stack([np.asarray(m, dtype=np.float64) for m in masks])
    nside = int(round(np.sqrt(M.shape[1] / 12)))
    n = M.shape[0]
    if n <= 8:
        C = ctx.mask_spectra(M, nside, lmax, niter)
        return {(i, j): C[i, j] for i in range(n) for j in range(n)}
    from .alm import Alm, alm2cl
    A = ctx.map2alm(M, nside, lmax, niter)
    alms = [Alm(lmax, lmax, a) for a in A]
    return {(i, j): alm2cl(alms[i], alms[j]) for i in range(n) for j in range(n)}
