"""HEALPix ring geometry and spherical-harmonic transforms.

The reference obtains mask spectra with Healpix.jl's `map2alm`
(src/modecoupling.jl:250-251, src/workspace.jl:155,163) and `alm2cl`.
`map2alm` / `alm2map` and the polarised `map2alm_pol` / `alm2map_pol` run on
the GPU (psx_map2alm / psx_alm2map / psx_map2alm_pol / psx_alm2map_pol:
rocFFT and direct ring DFTs, Legendre and Wigner-d sums in HIP) with
Healpix.jl's defaults (lmax = 3 nside - 1, three Jacobi refinement
iterations).

The `*_host` functions are direct numpy transforms of the same definitions
(FFT along each iso-latitude ring, recursion in l per m, vectorised over
rings), O(nside^3) and meant for small maps.  They are pinned by round trips,
Y_lm orthogonality, the Wigner d against sympy and the analytic quadrupole in
tests, and the GPU transforms are checked against them.
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


def alm2map_host(alm: Alm, nside: int) -> HealpixMap:
    """Synthesis on the host (numpy)."""
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


def map2alm_host(m, lmax: int | None = None, niter: int = 3) -> Alm:
    """Analysis on the host (numpy) with `niter` Jacobi refinements."""
    hm = m if isinstance(m, HealpixMap) else HealpixMap(m)
    lmax = nside2lmax(hm.nside) if lmax is None else int(lmax)
    info = RingInfo(hm.nside)
    alm = _map2alm_once(hm.pixels, info, lmax)
    for _ in range(niter):
        resid = hm.pixels - alm2map_host(alm, hm.nside).pixels
        alm = Alm(lmax, lmax, alm.alm + _map2alm_once(resid, info, lmax).alm)
    return alm


def map2alm(m, lmax: int | None = None, niter: int = 3, ctx=None) -> Alm:
    """Spherical-harmonic analysis on the GPU with `niter` Jacobi refinements
    (Healpix.jl's map2alm; default lmax = 3 nside - 1, niter = 3)."""
    from ._abi import default_context
    hm = m if isinstance(m, HealpixMap) else HealpixMap(m)
    lmax = nside2lmax(hm.nside) if lmax is None else int(lmax)
    ctx = ctx or default_context()
    return Alm(lmax, lmax, ctx.map2alm(hm.pixels, hm.nside, lmax, niter)[0])


def alm2map(alm: Alm, nside: int, ctx=None) -> HealpixMap:
    """Spherical-harmonic synthesis on the GPU (Healpix.jl's alm2map)."""
    from ._abi import default_context
    if alm.mmax != alm.lmax:
        raise ArgumentError("alm2map needs mmax = lmax")
    ctx = ctx or default_context()
    return HealpixMap(ctx.alm2map(alm.alm, nside, alm.lmax)[0], nside)


# ---- spin 2 (polarisation) -----------------------------------------------------
# Convention of HEALPix (and Healpix.jl): (Q +- iU)(n) = sum_lm _{+-2}a_lm
# _{+-2}Y_lm(n), a_E = -(_2a + _-2a)/2, a_B = i(_2a - _-2a)/2, with
# _sY_lm(theta, phi) = (-1)^s sqrt((2l+1)/(4 pi)) d^l_{m,-s}(theta) e^{i m phi}.
# E.g. a pure E mode a_E(2,0) = 1 gives Q = -sqrt(15/(32 pi)) sin^2 theta, U = 0.

def _d_rows(m: int, mp: int, lmax: int, cth: np.ndarray) -> np.ndarray:
    """Wigner d^l_{m,mp}(theta) for l = 0..lmax on every ring (zero for
    l < max(|m|, |mp|)), by the three-term recurrence in l; m >= 0, |mp| <= 2."""
    out = np.zeros((lmax + 1, cth.size))
    l0 = max(m, abs(mp))
    if l0 > lmax:
        return out
    c2 = np.sqrt(np.maximum(0.0, 0.5 * (1.0 + cth)))   # cos(theta/2)
    s2 = np.sqrt(np.maximum(0.0, 0.5 * (1.0 - cth)))   # sin(theta/2)
    lg = lambda x: float(np.sum(np.log(np.arange(1, x + 1)))) if x > 0 else 0.0

    def start(j, a, b):
        # d^j_{j,b} = (-1)^(j-b) sqrt((2j)! / ((j+b)! (j-b)!)) c^(j+b) s^(j-b)
        coef = 0.5 * (lg(2 * j) - lg(j + b) - lg(j - b))
        with np.errstate(divide="ignore"):
            v = coef + (j + b) * np.log(c2) + (j - b) * np.log(s2)
        return (-1.0) ** (j - b) * np.exp(v)

    if m >= abs(mp):
        d0 = start(m, m, mp)
    elif mp > 0:                       # d^l_{m,mp} = (-1)^(m-mp) d^l_{mp,m}
        d0 = (-1.0) ** (m - mp) * start(mp, mp, m)
    else:                              # d^l_{m,mp} = d^l_{-mp,-m}
        d0 = start(-mp, -mp, -m)
    out[l0] = d0
    prev = np.zeros_like(d0)
    cur = d0
    for l in range(l0, lmax):
        if l == 0:
            nxt = cth * cur            # m = mp = 0 is not needed for spin 2
        else:
            a = (2 * l + 1) * (l * (l + 1) * cth - m * mp)
            b = (l + 1) * np.sqrt(max(0.0, (l * l - m * m) * (l * l - mp * mp)))
            den = l * np.sqrt(((l + 1) ** 2 - m * m) * ((l + 1) ** 2 - mp * mp))
            nxt = (a * cur - b * prev) / den
        out[l + 1] = nxt
        prev, cur = cur, nxt
    return out


def _ring_phases_c(info: RingInfo, z: np.ndarray, mmax: int) -> np.ndarray:
    """F[k, m] = sum_j z_j e^{-i m phi_j} for a complex ring field z."""
    nr = info.first.size
    F = np.zeros((nr, mmax + 1), dtype=np.complex128)
    ms = np.arange(mmax + 1)
    for k in range(nr):
        n = int(info.count[k])
        ring = z[info.first[k]: info.first[k] + n]
        f = np.fft.fft(ring)
        F[k] = f[ms % n] * np.exp(-1j * ms * info.phi0[k])
    return F


def _pol_analysis_once(Q, U, info: RingInfo, lmax: int):
    w = 4 * np.pi / Q.size
    Fp = _ring_phases_c(info, Q + 1j * U, lmax)
    Fm = _ring_phases_c(info, Q - 1j * U, lmax)
    aE = np.zeros(num_alms(lmax), dtype=np.complex128)
    aB = np.zeros(num_alms(lmax), dtype=np.complex128)
    norm = np.sqrt((2 * np.arange(lmax + 1) + 1) / (4 * np.pi))
    for m in range(lmax + 1):
        dp = _d_rows(m, -2, lmax, info.cth) * norm[:, None]   # _2Y  (s = 2): d^l_{m,-2}
        dm = _d_rows(m, 2, lmax, info.cth) * norm[:, None]    # _-2Y (s = -2): d^l_{m,2}
        a2 = w * (dp @ Fp[:, m])
        am2 = w * (dm @ Fm[:, m])
        i0 = alm_index(lmax, m, m)
        sl = slice(i0, i0 + lmax - m + 1)
        aE[sl] = (-(a2 + am2) / 2)[m:]
        aB[sl] = (1j * (a2 - am2) / 2)[m:]
    for a in (aE, aB):   # l < 2 carry no spin-2 power
        for m in range(min(2, lmax + 1)):
            for l in range(m, min(2, lmax + 1)):
                a[alm_index(lmax, l, m)] = 0.0
    return aE, aB


def alm2map_pol_host(almT: Alm, almE: Alm, almB: Alm, nside: int):
    """(I, Q, U) maps from T, E, B coefficients (HEALPix convention), numpy."""
    info = RingInfo(nside)
    lmax = almE.lmax
    nr = info.first.size
    I = alm2map_host(almT, nside).pixels
    Gp = np.zeros((nr, lmax + 1), dtype=np.complex128)
    Gm = np.zeros((nr, lmax + 1), dtype=np.complex128)
    norm = np.sqrt((2 * np.arange(lmax + 1) + 1) / (4 * np.pi))
    for m in range(lmax + 1):
        aE = np.zeros(lmax + 1, dtype=np.complex128)
        aB = np.zeros(lmax + 1, dtype=np.complex128)
        aE[m:] = almE.m_slice(m)
        aB[m:] = almB.m_slice(m)
        a2 = -(aE + 1j * aB)
        am2 = -(aE - 1j * aB)
        dp = _d_rows(m, -2, lmax, info.cth) * norm[:, None]
        dm = _d_rows(m, 2, lmax, info.cth) * norm[:, None]
        Gp[:, m] = a2 @ dp
        Gm[:, m] = am2 @ dm
    P = np.zeros(nside2npix(nside), dtype=np.complex128)
    for k in range(nr):
        n = int(info.count[k])
        phi = info.phi0[k] + 2 * np.pi * np.arange(n) / n
        ms = np.arange(0, lmax + 1)
        E = np.exp(1j * np.outer(phi, ms))
        val = E @ Gp[k] + np.conj(E[:, 1:]) @ np.conj(Gm[k, 1:])
        P[info.first[k]: info.first[k] + n] = val
    return HealpixMap(I, nside), HealpixMap(P.real.copy(), nside), HealpixMap(P.imag.copy(), nside)


def map2alm_pol_host(I, Q, U, lmax: int | None = None, niter: int = 3):
    """(T, E, B) coefficients of an (I, Q, U) map with `niter` Jacobi
    refinements (Healpix.jl's polarised map2alm), numpy."""
    Im = I if isinstance(I, HealpixMap) else HealpixMap(I)
    Qp = np.asarray(Q.pixels if isinstance(Q, HealpixMap) else Q, dtype=np.float64)
    Up = np.asarray(U.pixels if isinstance(U, HealpixMap) else U, dtype=np.float64)
    ns = Im.nside
    lmax = nside2lmax(ns) if lmax is None else int(lmax)
    info = RingInfo(ns)
    T = map2alm_host(Im, lmax=lmax, niter=niter)
    aE, aB = _pol_analysis_once(Qp, Up, info, lmax)
    for _ in range(niter):
        _, q, u = alm2map_pol_host(Alm(lmax), Alm(lmax, lmax, aE), Alm(lmax, lmax, aB), ns)
        dE, dB = _pol_analysis_once(Qp - q.pixels, Up - u.pixels, info, lmax)
        aE = aE + dE
        aB = aB + dB
    return T, Alm(lmax, lmax, aE), Alm(lmax, lmax, aB)


def _pixels(m) -> np.ndarray:
    return np.asarray(m.pixels if isinstance(m, HealpixMap) else m, dtype=np.float64)


def map2alm_pol(I, Q, U, lmax: int | None = None, niter: int = 3, ctx=None):
    """(T, E, B) coefficients of an (I, Q, U) map on the GPU with `niter`
    Jacobi refinements (Healpix.jl's map2alm of a PolarizedHealpixMap, as
    master calls it, reference src/modecoupling.jl:328-329)."""
    from ._abi import default_context
    iqu = np.stack([_pixels(I), _pixels(Q), _pixels(U)])
    ns = HealpixMap(iqu[0]).nside
    lmax = nside2lmax(ns) if lmax is None else int(lmax)
    ctx = ctx or default_context()
    teb = ctx.map2alm_pol(iqu, ns, lmax, niter)[0]
    return tuple(Alm(lmax, lmax, teb[i]) for i in range(3))


def alm2map_pol(almT: Alm, almE: Alm, almB: Alm, nside: int, ctx=None):
    """(I, Q, U) maps from T, E, B coefficients on the GPU."""
    from ._abi import default_context
    lmax = almT.lmax
    for a in (almT, almE, almB):
        if a.lmax != lmax or a.mmax != lmax:
            raise ArgumentError("alm2map_pol needs three alms with mmax = lmax = lmax of T")
    ctx = ctx or default_context()
    iqu = ctx.alm2map_pol(np.stack([almT.alm, almE.alm, almB.alm]), nside, lmax)[0]
    return tuple(HealpixMap(iqu[i], nside) for i in range(3))
