"""QuickPol beam matrices (reference src/beam.jl).

`quickpolXi` fills a banded Ξ matrix on the GPU (psx_quickpol_xi: one lane per
band cell, two Schulten-Gordon Wigner-3j families per lane, accumulated on the
fly); `quickpolW` is the scan-pattern spectrum it contracts, and `ku` the
spin weight k_u.  The banded matrix is held in BandedMatrices.jl's band
storage, as the reference's `SpectralArray` of a `BandedMatrix`.
"""
from __future__ import annotations

import numpy as np

from ._abi import ArgumentError, AxisError, Context, default_context
from .alm import Alm, alm_index
from .spectralarray import SpectralVector


class BandedMatrix:
    """Square n x n matrix with lower bandwidth l and upper bandwidth u in
    band storage: element (i, j), -l <= j - i <= u, at data[u + i - j, j]
    (0-based; BandedMatrices.jl's layout).  Indices here are 0-based, i.e. the
    multipoles 0..n-1 of the reference's SpectralArray axes."""

    __slots__ = ("n", "l", "u", "data")

    def __init__(self, n: int, l: int, u: int, data=None):
        if n < 1 or l < 0 or u < 0:
            raise ArgumentError("BandedMatrix needs n >= 1 and bandwidths >= 0")
        self.n, self.l, self.u = int(n), int(l), int(u)
        if data is None:
            data = np.zeros((self.l + self.u + 1, self.n), order="F")
        data = np.asarray(data, dtype=np.float64)
        if data.shape != (self.l + self.u + 1, self.n):
            raise AxisError("band storage must be (l + u + 1, n)")
        self.data = np.asfortranarray(data)

    @property
    def shape(self):
        return (self.n, self.n)

    def inband(self, i: int, j: int) -> bool:
        return 0 <= i < self.n and 0 <= j < self.n and -self.l <= j - i <= self.u

    def __getitem__(self, ij):
        i, j = ij
        return self.data[self.u + i - j, j] if self.inband(i, j) else 0.0

    def __setitem__(self, ij, v):
        i, j = ij
        if not self.inband(i, j):
            raise ArgumentError(f"({i}, {j}) is outside the band")
        self.data[self.u + i - j, j] = v

    def rowrange(self, i: int) -> range:
        """Columns of row i inside the band (BandedMatrices rowstart:rowstop)."""
        return range(max(0, i - self.l), min(self.n - 1, i + self.u) + 1)

    def todense(self) -> np.ndarray:
        A = np.zeros((self.n, self.n))
        for i in range(self.n):
            for j in self.rowrange(i):
                A[i, j] = self.data[self.u + i - j, j]
        return A


def quickpolW(alm1: Alm, alm2: Alm):
    """W_l = Re(a1_l0 a2_l0*) + 2 sum_{m=1}^{min(l, mmax)} Re(a1_lm a2_lm*),
    l = 0..min(lmax1, lmax2) (src/beam.jl:41-54); both sets are indexed with
    alm1's layout, as the reference does."""
    mmax = min(alm1.mmax, alm2.mmax)
    lmax = min(alm1.lmax, alm2.lmax)
    cl = np.zeros(lmax + 1)
    for m in range(0, mmax + 1):
        ls = np.arange(m, lmax + 1)
        if ls.size == 0:
            continue
        idx = alm_index(alm1.lmax, ls, m)
        if idx[-1] >= alm2.alm.size:
            raise ArgumentError("alm2 is shorter than alm1's layout requires")
        x = np.real(alm1.alm[idx] * np.conj(alm2.alm[idx]))
        cl[m:] += x if m == 0 else 2.0 * x
    return SpectralVector(cl)


def quickpolXi(Xi: BandedMatrix, nu1: int, nu2: int, s1: int, s2: int, w1: Alm, w2: Alm,
               ctx: Context | None = None) -> BandedMatrix:
    """quickpolΞ!(𝚵, ν₁, ν₂, s₁, s₂, ω₁, ω₂) (src/beam.jl:72-109), in place:
    Ξ[l'', l] = (-1)^(s1+s2+nu1+nu2) sum_{l'} W_{l'} (l' l l''; s1+nu1, -s1, -nu1)
    (l' l l''; s2+nu2, -s2, -nu2) for l'' >= 2 and band columns l >= 2, with
    W = quickpolW(w1, w2); every other band entry is multiplied by the sign.
    Terms with l' beyond W's length count as zero."""
    if not isinstance(Xi, BandedMatrix):
        raise ArgumentError("Ξ must be a BandedMatrix")
    W = quickpolW(w1, w2).parent
    return quickpolXi_from_W(Xi, nu1, nu2, s1, s2, W, ctx=ctx)


def quickpolXi_from_W(Xi: BandedMatrix, nu1: int, nu2: int, s1: int, s2: int, W,
                      ctx: Context | None = None) -> BandedMatrix:
    """The same with the scan spectrum W (0-based) given directly."""
    ctx = ctx or default_context()
    lmax = Xi.n - 1
    ctx.quickpol_xi(nu1, nu2, s1, s2, lmax, Xi.l, Xi.u, np.asarray(W, dtype=np.float64),
                    band=Xi.data)
    return Xi


def ku(u: int, dtype=float):
    """k_u: 1 for u = 0, 1/2 for |u| = 2 (src/beam.jl:116-125)."""
    if u == 0:
        return dtype(1)
    if abs(u) == 2:
        return dtype(0.5)
    raise ArgumentError("Defined only for u ∈ {-2, 0, 2}.")
