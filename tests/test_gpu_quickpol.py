"""GPU: QuickPol Ξ (psx_quickpol_xi, reference src/beam.jl:72-101) against the
oracle, which is pinned by exact sympy sums and the TT mode-coupling reduction
on the CPU.  Band cells are compared elementwise relative to the largest
entry of the band (a family product sum can cancel to near zero)."""
import numpy as np
import pytest
from helpers import mask_like_W

pytestmark = pytest.mark.gpu

SPINS = [(0, 0, 0, 0), (0, 0, 2, 2), (1, -2, 2, 0), (2, 2, -2, 2), (-4, 3, 2, -2), (0, -2, 2, 0)]


def _cmp(got, want):
    scale = np.abs(want).max()
    return float(np.abs(got - want).max() / scale) if scale > 0 else float(np.abs(got).max())


@pytest.mark.parametrize("nu1,nu2,s1,s2", SPINS)
def test_quickpol_xi_vs_oracle(ctx, oracle, nu1, nu2, s1, s2):
    lmax, bl, bu = 300, 6, 9
    rng = np.random.default_rng(7)
    W = rng.uniform(-1.0, 1.0, 2 * lmax + 1)
    start = np.asfortranarray(rng.uniform(1.0, 2.0, (bl + bu + 1, lmax + 1)))
    got = ctx.quickpol_xi(nu1, nu2, s1, s2, lmax, bl, bu, W, band=start.copy(order="F"))
    want = oracle.quickpol_xi(nu1, nu2, s1, s2, lmax, bl, bu, W, band=start.copy(order="F"))
    assert _cmp(got, want) <= 1e-12
    # rows and columns 0, 1 are only multiplied by the sign
    sgn = (-1) ** ((s1 + s2 + nu1 + nu2) % 2)
    for lpp in range(0, 2):
        for l in range(max(0, lpp - bl), min(lmax, lpp + bu) + 1):
            assert got[bu + lpp - l, l] == sgn * start[bu + lpp - l, l]


def test_quickpol_xi_short_W_and_invalid_families(ctx, oracle):
    """W of length lmax + 1 (sums stop there) and nu = 4 (families with
    l'' < 4 have no valid 3j symbol and are zero)."""
    lmax, bl, bu = 120, 3, 3
    W = mask_like_W(lmax + 1, seed=2)
    for spins in ((4, 0, 2, 2), (0, 0, 0, 0)):
        got = ctx.quickpol_xi(*spins, lmax, bl, bu, W)
        want = oracle.quickpol_xi(*spins, lmax, bl, bu, W)
        assert _cmp(got, want) <= 1e-12
    got = ctx.quickpol_xi(4, 0, 2, 2, lmax, bl, bu, W)
    for lpp in (2, 3):
        for l in range(max(2, lpp - bl), lpp + bu + 1):
            assert got[bu + lpp - l, l] == 0.0


def test_quickpol_xi_lmax3071(ctx, oracle):
    """A full band of 21 diagonals at lmax 3071 (families of up to 6143 terms)."""
    lmax, bl, bu = 3071, 10, 10
    W = mask_like_W(2 * lmax + 1, seed=4)
    for spins in ((2, -2, 2, 2), (0, 0, 0, 2)):
        got = ctx.quickpol_xi(*spins, lmax, bl, bu, W)
        want = oracle.quickpol_xi(*spins, lmax, bl, bu, W)
        assert _cmp(got, want) <= 1e-12, spins


def test_quickpolXi_mirror(pkg, ctx, oracle):
    """quickpolΞ!(𝚵, ν₁, ν₂, s₁, s₂, ω₁, ω₂) through the Python mirror: W from
    quickpolW of two scan-weight alms, Ξ in a BandedMatrix."""
    from powerspectra_amd.beam import BandedMatrix, quickpolW, quickpolXi
    rng = np.random.default_rng(1)
    lmax_w = 161
    w1, w2 = pkg.Alm(lmax_w), pkg.Alm(lmax_w)
    for a in (w1, w2):
        a.alm[:] = rng.standard_normal(a.alm.size) + 1j * rng.standard_normal(a.alm.size)
    Xi = BandedMatrix(81, 4, 4)
    quickpolXi(Xi, 2, 0, 2, -2, w1, w2, ctx=ctx)
    want = oracle.quickpol_xi(2, 0, 2, -2, 80, 4, 4, quickpolW(w1, w2).parent)
    assert _cmp(Xi.data, want) <= 1e-12


@pytest.mark.parametrize("lmax,bl,bu", [(0, 0, 0), (1, 1, 0), (2, 0, 0), (3, 1, 2), (5, 5, 5)])
def test_quickpol_xi_small(ctx, oracle, lmax, bl, bu):
    """Tiny matrices: single diagonals, bands wider than the matrix, lmax < 2
    (only the sign is applied)."""
    rng = np.random.default_rng(lmax + 10 * bl)
    W = rng.uniform(-1.0, 1.0, 2 * lmax + 1)
    start = np.asfortranarray(rng.uniform(1.0, 2.0, (bl + bu + 1, lmax + 1)))
    for spins in ((0, 0, 0, 0), (1, 0, 2, 0)):
        got = ctx.quickpol_xi(*spins, lmax, bl, bu, W, band=start.copy(order="F"))
        want = oracle.quickpol_xi(*spins, lmax, bl, bu, W, band=start.copy(order="F"))
        # cells of the storage outside the matrix are untouched by both
        assert np.abs(got - want).max() <= 1e-13 * max(1.0, np.abs(want).max())
