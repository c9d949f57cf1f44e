"""Spherical-harmonic coefficient container and cross-spectra (alm2cl).

Layout follows Healpix.jl's `Alm` (used by the reference at
src/modecoupling.jl:192-197 and src/workspace.jl:202): complex coefficients
in m-major order, index(l, m) = m (2 lmax + 1 - m) / 2 + l for 0 <= m <= l.
"""
from __future__ import annotations

import numpy as np

from ._abi import ArgumentError


def num_alms(lmax: int, mmax: int | None = None) -> int:
    mmax = lmax if mmax is None else mmax
    return (mmax + 1) * (2 * lmax + 2 - mmax) // 2


def alm_index(lmax: int, l, m):
    """0-based position of a_lm (Healpix.jl almIndex minus one)."""
    return m * (2 * lmax + 1 - m) // 2 + l


class Alm:
    """a_lm for 0 <= m <= l <= lmax (mmax = lmax)."""

    __slots__ = ("lmax", "mmax", "alm")

    def __init__(self, lmax: int, mmax: int | None = None, alm=None):
        self.lmax = int(lmax)
        self.mmax = self.lmax if mmax is None else int(mmax)
        n = num_alms(self.lmax, self.mmax)
        if alm is None:
            alm = np.zeros(n, dtype=np.complex128)
        alm = np.asarray(alm, dtype=np.complex128)
        if alm.shape != (n,):
            raise ArgumentError(f"expected {n} coefficients for lmax={lmax}, mmax={self.mmax}")
        self.alm = alm

    def __getitem__(self, lm):
        l, m = lm
        return self.alm[alm_index(self.lmax, l, m)]

    def __setitem__(self, lm, v):
        l, m = lm
        self.alm[alm_index(self.lmax, l, m)] = v

    def m_slice(self, m: int) -> np.ndarray:
        """Coefficients a_{l m} for l = m .. lmax (contiguous)."""
        i0 = alm_index(self.lmax, m, m)
        return self.alm[i0: i0 + self.lmax - m + 1]


def alm2cl(a: Alm, b: Alm | None = None) -> np.ndarray:
    """C_l = (2l+1)^-1 [Re(a_l0 b_l0*) + 2 sum_{m>=1} Re(a_lm b_lm*)], l = 0..lmax."""
    b = a if b is None else b
    if a.lmax != b.lmax or a.mmax != b.mmax:
        raise ArgumentError("alm2cl needs coefficient sets with equal lmax and mmax")
    lmax = a.lmax
    cl = np.zeros(lmax + 1)
    for m in range(0, a.mmax + 1):
        x = np.real(a.m_slice(m) * np.conj(b.m_slice(m)))
        cl[m:] += x if m == 0 else 2.0 * x
    return cl / (2.0 * np.arange(lmax + 1) + 1.0)
