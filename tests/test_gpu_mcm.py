"""GPU parity: mode-coupling matrices through the C ABI vs the CPU oracle.

Reference: inner_mcm⁰⁰!/⁰²!/⁺⁺!/⁻⁻! (src/modecoupling.jl:78-159) and the Ξ
projectors (:3-66).  Tolerance: RTOL_FP64 = 1e-12 relative (elementwise over
|ref| >= 1e-6 max|ref|, and normwise).
"""
import ctypes

import numpy as np
import pytest
from helpers import RTOL_FP64, mask_like_W, norm_err, rel_err

pytestmark = pytest.mark.gpu

KINDS = [0, 1, 2, 3]


def _cmp(ctx, oracle, kind, lmin, lmax, W):
    M = ctx.mcm(kind, lmin, lmax, W)
    R = oracle.mcm(kind, lmin, lmax, W)
    assert np.isfinite(M).all()
    e, ne = rel_err(M, R), norm_err(M, R)
    assert e <= RTOL_FP64 and ne <= RTOL_FP64, (kind, lmin, lmax, W.size, e, ne)
    return M


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("lmin,lmax", [(0, 0), (0, 1), (0, 2), (0, 3), (1, 3), (0, 17), (2, 17),
                                       (0, 64), (5, 64), (0, 130), (63, 200)])
def test_small_vs_oracle(ctx, oracle, kind, lmin, lmax):
    for LW in sorted({1, 7, lmax + 1, 2 * lmax + 1}):
        W = mask_like_W(LW, seed=lmax + LW, cross=(kind == 1))
        _cmp(ctx, oracle, kind, lmin, lmax, W)


@pytest.mark.parametrize("kind", KINDS)
def test_medium_vs_oracle(ctx, oracle, kind):
    lmax = 400
    W = mask_like_W(lmax + 1, seed=11)
    _cmp(ctx, oracle, kind, 0, lmax, W)


def test_pm_and_set_match_single_blocks(ctx):
    lmax = 333
    LW = lmax + 1
    Wt, Wte, Wet, Wp = (mask_like_W(LW, seed=s, cross=(s in (2, 3))) for s in (1, 2, 3, 4))
    singles = [ctx.mcm(0, 0, lmax, Wt), ctx.mcm(1, 0, lmax, Wte), ctx.mcm(1, 0, lmax, Wet),
               ctx.mcm(2, 0, lmax, Wp), ctx.mcm(3, 0, lmax, Wp)]
    P, Mm = ctx.mcm_pm(0, lmax, Wp)
    assert rel_err(P, singles[3]) <= 1e-15 and rel_err(Mm, singles[4]) <= 1e-15
    outs = ctx.mcm_set(0, lmax, Wt, Wte, Wet, Wp)
    for o, s in zip(outs, singles):
        assert rel_err(o, s) <= 1e-15
    # partial set (TE only) goes through the single-block path
    outs2 = ctx.mcm_set(0, lmax, None, Wte, None, None)
    assert outs2[0] is None and rel_err(outs2[1], singles[1]) == 0.0


def test_full_sky_identity(ctx):
    """Full sky: W_0 = 4 pi, W_l>0 = 0 -> M00 = I, M02 = M++ = I (l >= 2), M-- = 0."""
    lmax = 3071
    W = np.zeros(lmax + 1)
    W[0] = 4 * np.pi
    n = lmax + 1
    I = np.eye(n)
    I2 = np.diag((np.arange(n) >= 2).astype(float))
    M00 = ctx.mcm(0, 0, lmax, W)
    assert np.abs(M00 - I).max() < 1e-13
    M02 = ctx.mcm(1, 0, lmax, W)
    assert np.abs(M02 - I2).max() < 1e-13
    P, Mm = ctx.mcm_pm(0, lmax, W)
    assert np.abs(P - I2).max() < 1e-13
    assert np.abs(Mm).max() == 0.0


def test_flat_W_sum_rule(ctx):
    """W_l = 4 pi c for l <= 2 lmax: the 3j sum rule gives M00[l1,l2] = c (2l2+1)
    and M++ + M-- = c (2l2+1) for l1, l2 >= 2 (full-size lmax 3071)."""
    lmax = 3071
    c = 0.37
    W = np.full(2 * lmax + 1, 4 * np.pi * c)
    n = lmax + 1
    target = c * (2 * np.arange(n) + 1.0)[None, :] * np.ones((n, 1))
    M00 = ctx.mcm(0, 0, lmax, W)
    assert rel_err(M00, target) < 1e-13
    P, Mm = ctx.mcm_pm(0, lmax, W)
    assert rel_err((P + Mm)[2:, 2:], target[2:, 2:]) < 1e-13


@pytest.mark.parametrize("kind", KINDS)
def test_full_size_rows_vs_oracle(ctx, oracle, kind):
    """lmax 3071 (the benchmark size), W of length lmax+1 as in the reference:
    a sample of rows against the oracle, plus the (2l2+1)/(2l1+1) symmetry."""
    lmax = 3071
    W = mask_like_W(lmax + 1, seed=7, cross=(kind == 1))
    M = ctx.mcm(kind, 0, lmax, W)
    assert np.isfinite(M).all()
    R = np.full((lmax + 1, lmax + 1), np.nan, order="F")
    for lo, hi in ((0, 3), (1021, 1023), (2900, 2901), (3070, 3071)):
        oracle.mcm_rows(kind, 0, lmax, lo, hi, W, out=R)
    sel = np.isfinite(R)
    e = rel_err(np.where(sel, M, np.nan), R)
    assert e <= RTOL_FP64, e
    ell = 2.0 * np.arange(lmax + 1) + 1.0
    X = M / ell[None, :]
    assert rel_err(X, X.T) < 1e-15


def test_lmin_block_is_subblock(ctx):
    lmax = 250
    W = mask_like_W(lmax + 1, seed=3)
    full = ctx.mcm(2, 0, lmax, W)
    sub = ctx.mcm(2, 40, lmax, W)
    assert np.array_equal(sub, full[40:, 40:])


def test_errors(ctx, pkg):
    W = mask_like_W(50)
    with pytest.raises(pkg.ArgumentError):
        ctx.mcm(7, 0, 40, W)
    with pytest.raises(pkg.ArgumentError):
        ctx.mcm(0, 10, 5, W)
    L = ctx.lib
    out = np.zeros((41, 41), order="F")
    st = L.psx_mcm(ctx.handle, 0, 0, 40, W.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                   W.size, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 30)
    assert st == 2  # PSX_ERR_AXES


def test_reference_namaster_diagonals_gpu(ctx):
    """The HIP path against the reference's NaMaster fixtures
    (test/data/mcm_{TT,EE,TE}_diag.txt): with the even-l3 mask spectrum fitted
    from the TT diagonal (oracle families, CPU), the GPU diagonals reproduce
    TT and predict EE (M⁺⁺) and TE."""
    from test_oracle import _namaster_fit
    W, tt, ee, te, _ = _namaster_fit()
    lmax = 767
    dtt = np.diag(ctx.mcm(0, 0, lmax, W))[2:767]
    dee = np.diag(ctx.mcm(2, 0, lmax, W))[2:767]
    dte = np.diag(ctx.mcm(1, 0, lmax, W))[2:767]
    assert np.max(np.abs(dtt - tt) / np.abs(tt)) < 1e-12
    assert np.max(np.abs(dee - ee) / np.abs(ee)) < 5e-7
    assert np.max(np.abs(dte - te) / np.abs(te)) < 5e-7
