"""GPU parity: coupled covariance blocks through the C ABI vs the CPU oracle.

Reference: loop_covTTTT!/EEEE!/TTTE!/TETE!/TEEE!/TEEE_planck!/TTEE!
(src/covariance.jl:92-446).  Tolerance RTOL_FP64 = 1e-12.
"""
import numpy as np
import pytest
from helpers import RTOL_FP64, mask_like_W, norm_err, rel_err, spectra_set

pytestmark = pytest.mark.gpu

ARITY = {0: (8, 4, 4), 1: (8, 4, 4), 2: (4, 4, 2), 3: (5, 4, 2), 4: (4, 4, 2), 5: (4, 4, 2),
         6: (2, 4, 0)}


def _inputs(kind, lmax, LW, seed=0):
    nW, nS, nR = ARITY[kind]
    spec, rat = spectra_set(lmax, seed)
    if kind in (2, 3, 4, 5, 6):  # TE spectra may be negative (not under a sqrt)
        rng = np.random.default_rng(seed + 100)
        for i in (2, 3):
            spec[i] = spec[i] * np.where(rng.uniform(size=lmax + 1) < 0.3, -1.0, 1.0)
        if kind == 6:
            for i in (0, 1):
                spec[i] = spec[i] * np.where(rng.uniform(size=lmax + 1) < 0.3, -1.0, 1.0)
    W = [mask_like_W(LW, seed=seed * 10 + k, scale=10.0 ** -k) for k in range(nW)]
    return spec[:nS], rat[:nR], W


@pytest.mark.parametrize("kind", range(7))
@pytest.mark.parametrize("lmin,lmax", [(0, 0), (0, 3), (2, 40), (0, 180)])
def test_cov_vs_oracle(ctx, oracle, kind, lmin, lmax):
    for LW in sorted({lmax + 1, 2 * lmax + 1}):
        spec, rat, W = _inputs(kind, lmax, LW, seed=lmax + kind)
        C = ctx.cov(kind, lmin, lmax, spec, rat, W)
        R = oracle.cov(kind, lmin, lmax, spec, rat, W)
        assert np.isfinite(C).all()
        assert np.array_equal(C, C.T)
        e, ne = rel_err(C, R), norm_err(C, R)
        assert e <= RTOL_FP64 and ne <= RTOL_FP64, (kind, lmin, lmax, LW, e, ne)


@pytest.mark.parametrize("kind", [0, 3])
def test_cov_full_size_rows(ctx, oracle, kind):
    """lmax 2508 (the covariance benchmark size): rows sampled against the oracle."""
    lmax = 2508
    spec, rat, W = _inputs(kind, lmax, lmax + 1, seed=5)
    C = ctx.cov(kind, 0, lmax, spec, rat, W)
    R = oracle.cov(kind, 0, lmax, spec, rat, W, rows=(2400, 2402))
    R2 = oracle.cov(kind, 0, lmax, spec, rat, W, rows=(3, 5))
    R = np.where(np.isfinite(R), R, R2)
    assert rel_err(np.where(np.isfinite(R), C, np.nan), R) <= RTOL_FP64


def test_full_sky_knox(ctx):
    """Full sky: every W_k = 4 pi w_k delta_{l,0} makes Ξ_k = w_k delta_{l1 l2} / (2 l1 + 1),
    so TTTT is diagonal with the reference's coefficient algebra
    (src/covariance.jl:110-118) applied to w_k."""
    lmax = 300
    spec, rat = spectra_set(lmax, 3)
    w = np.array([1.0, 0.8, 0.3, 0.25, 0.2, 0.15, 0.05, 0.04])
    W = []
    for k in range(8):
        x = np.zeros(lmax + 1)
        x[0] = 4 * np.pi * w[k]
        W.append(x)
    C = ctx.cov(0, 0, lmax, spec, rat, W)
    ip, jq, iq, jp = spec
    rip, rjq, riq, rjp = rat
    l = np.arange(lmax + 1)
    diag = (np.sqrt(ip * ip * jq * jq) * w[0] + np.sqrt(iq * iq * jp * jp) * w[1]
            + np.sqrt(ip * ip) * w[2] * rjq * rjq + np.sqrt(jq * jq) * w[3] * rip * rip
            + np.sqrt(iq * iq) * w[4] * rjp * rjp + np.sqrt(jp * jp) * w[5] * riq * riq
            + w[6] * rip * rjq * rip * rjq + w[7] * riq * rjp * riq * rjp) / (2 * l + 1)
    assert rel_err(np.diag(C), diag) < 1e-14
    off = C - np.diag(np.diag(C))
    assert np.abs(off).max() == 0.0


def test_domain_error(ctx, pkg):
    lmax = 30
    spec, rat, W = _inputs(0, lmax, lmax + 1)
    spec[0] = spec[0].copy()
    spec[0][5] = -1.0
    with pytest.raises(pkg.DomainError):
        ctx.cov(0, 0, lmax, spec, rat, W)


def test_bad_arity(ctx, pkg):
    lmax = 10
    spec, rat, W = _inputs(0, lmax, lmax + 1)
    with pytest.raises(pkg.ArgumentError):
        ctx.cov(0, 0, lmax, spec, rat, W[:7])
