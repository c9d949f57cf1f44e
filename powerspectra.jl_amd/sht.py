"""HEALPix ring geometry and spherical-harmonic transforms (numpy, host).

Upstream of the GPU path: the reference obtains mask spectra with
Healpix.jl's `map2alm` (src/modecoupling.jl:250-251, src/workspace.jl:155,163)
and `alm2cl`.  This is a direct ring-by-ring transform -- FFT along each
iso-latitude ring, associated-Legendre recursion in l per m, vectorised over
rings -- with the same defaults as Healpix.jl (lmax = 3 nside - 1, three
Jacobi refinement iterations).  It is O(nside^3) on the host and meant for
nside <= 512; pinned by round trips and Y_lm orthogonality in tests.
"""
from __future__ import annotations

import numpy as np

from ._abi import ArgumentError
from .alm import Alm, alm_index, num_alms


class HealpixMap:
    """A ring-ordered HEALPix map (pixels[0:12 nside^2])."""

    __slots__ = ("pixels", "nside")

    def __init__(self, pixels, nside: int | None = None):
        p = np.asarray(pixels, dtype=np.float64)
        npix = p.size
        ns = int(round(np.sqrt(npix / 12))) if nside is None else int(nside)
        if 12 * ns * ns != npix:
            raise ArgumentError(f"{npix} pixels is not a HEALPix map")
        self.pixels = p
        self.nside = ns

    @property
    def npix(self) -> int:
        return self.pixels.size

    def __array__(self, dtype=None, copy=None):
        return np.asarray(self.pixels, dtype=dtype)


def nside2npix(nside: int) -> int:
    return 12 * nside * nside


def nside2lmax(nside: int) -> int:
    """3 nside - 1 (reference src/util.jl:47)."""
    return 3 * nside - 1


def pixsize(m: HealpixMap) -> float:
    """4 pi / Npix (reference src/workspace.jl:138)."""
    return 4 * np.pi / m.npix


class RingInfo:
    """Per-ring geometry: first pixel, pixel count, cos/sin theta, phi0."""

    def __init__(self, nside: int):
        ns = nside
        nr = 4 * ns - 1
        first = np.zeros(nr, dtype=np.int64)
        count = np.zeros(nr, dtype=np.int64)
        z = np.zeros(nr)
        phi0 = np.zeros(nr)
        for k in range(nr):
            i = k + 1
            if i < ns:
                count[k] = 4 * i
                first[k] = 2 * i * (i - 1)
                z[k] = 1.0 - i * i / (3.0 * ns * ns)
                phi0[k] = np.pi / (4.0 * i)
            elif i <= 3 * ns:
                count[k] = 4 * ns
                first[k] = 2 * ns * (ns - 1) + (i - ns) * 4 * ns
                z[k] = 4.0 / 3.0 - 2.0 * i / (3.0 * ns)
                s = (i - ns + 1) % 2
                phi0[k] = np.pi / (4.0 * ns) * s
            else:
                ii = 4 * ns - i
                count[k] = 4 * ii
                first[k] = 12 * ns * ns - 2 * ii * (ii + 1)
                z[k] = -(1.0 - ii * ii / (3.0 * ns * ns))
                phi0[k] = np.pi / (4.0 * ii)
        self.nside = ns
        self.first, self.count, self.cth, self.phi0 = first, count, z, phi0
        self.sth = np.sqrt(np.maximum(0.0, (1.0 - z) * (1.0 + z)))


def _lambda_lm_rows(m: int, lmax: int, cth: np.ndarray, sth: np.ndarray) -> np.ndarray:
    """lambda_lm(theta) for l = m..lmax on every ring (orthonormal, Y_lm = lambda e^{i m phi})."""
    nl = lmax - m + 1
    out = np.zeros((nl, cth.size))
    # lambda_mm = (-1)^m sqrt((2m+1)/(4 pi) prod_{k<=m} (2k-1)/(2k)) sin^m
    logc = 0.5 * (np.log((2 * m + 1) / (4 * np.pi)) + np.sum(np.log((2.0 * np.arange(1, m + 1) - 1)
                                                                   / (2.0 * np.arange(1, m + 1)))))
    with np.errstate(divide="ignore"):
        logs = m * np.log(sth)
    lam = np.where(sth > 0, np.exp(logc + logs), 1.0 if m == 0 else 0.0) * (-1.0) ** m
    out[0] = lam
    if nl > 1:
        out[1] = cth * np.sqrt(2 * m + 3.0) * lam
    for j in range(2, nl):
        l = m + j
        a = np.sqrt((4.0 * l * l - 1.0) / (l * l - m * m))
        b = np.sqrt(((l - 1.0) ** 2 - m * m) / (4.0 * (l - 1.0) ** 2 - 1.0))
        out[j] = a * (cth * out[j - 1] - b * out[j - 2])
    return out


def _ring_phases(info: RingInfo, pixels: np.ndarray, mmax: int) -> np.ndarray:
    """F[k, m] = sum_j pix e^{-i m phi_j} on ring k, m = 0..mmax."""
    nr = info.first.size
    F = np.zeros((nr, mmax + 1), dtype=np.complex128)
    ms = np.arange(mmax + 1)
    for k in range(nr):
        n = int(info.count[k])
        ring = pixels[info.first[k]: info.first[k] + n]
        f = np.fft.fft(ring)  # sum_j x_j e^{-2 pi i j q / n}
        F[k] = f[ms % n] * np.exp(-1j * ms * info.phi0[k])
    return F


def _map2alm_once(pixels: np.ndarray, info: RingInfo, lmax: int) -> Alm:
    F = _ring_phases(info, pixels, lmax)
    w = 4 * np.pi / pixels.size
    out = np.zeros(num_alms(lmax), dtype=np.complex128)
    for m in range(lmax + 1):
        lam = _lambda_lm_rows(m, lmax, info.cth, info.sth)
        i0 = alm_index(lmax, m, m)
        out[i0: i0 + lmax - m + 1] = w * (lam @ F[:, m])
    return Alm(lmax, lmax, out)


def alm2map(alm: Alm, nside: int) -> HealpixMap:
    info = RingInfo(nside)
    lmax = alm.lmax
    nr = info.first.size
    G = np.zeros((nr, lmax + 1), dtype=np.complex128)
    for m in range(lmax + 1):
        lam = _lambda_lm_rows(m, lmax, info.cth, info.sth)
        G[:, m] = alm.m_slice(m) @ lam
    pixels = np.zeros(nside2npix(nside))
    for k in range(nr):
        n = int(info.count[k])
        phi = info.phi0[k] + 2 * np.pi * np.arange(n) / n
        ms = np.arange(1, lmax + 1)
        val = np.real(G[k, 0]) * np.ones(n)
        if lmax >= 1:
            val += 2.0 * np.real(np.exp(1j * np.outer(phi, ms)) @ G[k, 1:])
        pixels[info.first[k]: info.first[k] + n] = val
    return HealpixMap(pixels, nside)


def map2alm(m, lmax: int | None = None, niter: int = 3) -> Alm:
    """Spherical-harmonic analysis with `niter` Jacobi refinements (Healpix.jl default 3)."""
    hm = m if isinstance(m, HealpixMap) else HealpixMap(m)
    lmax = nside2lmax(hm.nside) if lmax is None else int(lmax)
    info = RingInfo(hm.nside)
    alm = _map2alm_once(hm.pixels, info, lmax)
    for _ in range(niter):
        resid = hm.pixels - alm2map(alm, hm.nside).pixels
        alm = Alm(lmax, lmax, alm.alm + _map2alm_once(resid, info, lmax).alm)
    return alm
