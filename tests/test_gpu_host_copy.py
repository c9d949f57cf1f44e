"""GPU: the host-pointer entry points against the device-pointer ones.

A host-pointer call runs large contractions in l1 bands and copies the output
columns each band completes to host memory while the later bands compute.
Same kernel and same pairs, so the host results must equal the device
results bit for bit.  Each case first runs a call with other inputs so that
the context's transfer buffer holds different values, and fills the host
outputs with NaN: a column copied before its band had finished, or never
copied, shows up as a mismatch.
"""
import ctypes

import numpy as np
import pytest
from helpers import mask_like_W, spectra_set

pytestmark = pytest.mark.gpu

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _nan(rows, cols):
    return np.full((rows, cols), np.nan, order="F")


class _Dev:
    """Device buffers freed on exit."""

    def __init__(self, ctx):
        self.ctx, self.ps = ctx, []

    def put(self, a):
        p = self.ctx.alloc(a.nbytes)
        self.ps.append(p)
        self.ctx.h2d(p, a)
        return p

    def empty(self, nbytes):
        p = self.ctx.alloc(nbytes)
        self.ps.append(p)
        return p

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        for p in self.ps:
            self.ctx.free(p)


def _set_device(ctx, lmin, lmax, Ws):
    n = lmax - lmin + 1
    with _Dev(ctx) as d:
        dW = [d.put(w) for w in Ws]
        dM = [d.empty(8 * n * n) for _ in range(5)]
        ctx.check(ctx.lib.psx_mcm_set_d(ctx.handle, lmin, lmax, Ws[0].size, *dW, *dM, n),
                  "psx_mcm_set_d")
        outs = [np.empty((n, n), order="F") for _ in range(5)]
        for o, p in zip(outs, dM):
            ctx.d2h(o, p)
    return outs


def _set_host(ctx, lmin, lmax, Ws, ld):
    n = lmax - lmin + 1
    outs = [_nan(ld, n) for _ in range(5)]
    ctx.check(ctx.lib.psx_mcm_set(ctx.handle, lmin, lmax, Ws[0].size, *[_ptr(w) for w in Ws],
                                  *[_ptr(o) for o in outs], ld), "psx_mcm_set")
    return outs


@pytest.mark.parametrize("lmin,lmax,pad", [(0, 3071, 0), (2, 1500, 7), (0, 950, 0), (0, 500, 0),
                                          (4, 130, 3)])
def test_set_host_equals_device(pkg, lmin, lmax, pad):
    ctx = pkg.Context(0)
    try:
        LW = lmax + 1
        Ws = [mask_like_W(LW, seed=s, cross=(s in (2, 3))) for s in (1, 2, 3, 4)]
        other = [2.0 * w for w in Ws]
        n = lmax - lmin + 1
        _set_host(ctx, lmin, lmax, other, n + pad)  # leaves other values in the buffers
        ref = _set_device(ctx, lmin, lmax, Ws)
        got = _set_host(ctx, lmin, lmax, Ws, n + pad)
        for k in range(5):
            assert np.array_equal(got[k][:n], ref[k]), (k, lmin, lmax, pad)
            if pad:  # rows below the block are left alone
                assert np.isnan(got[k][n:]).all()
    finally:
        ctx.close()


@pytest.mark.parametrize("lmin,lmax", [(3, 2600), (0, 767), (1, 40)])
@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_single_block_host_equals_device(pkg, kind, lmin, lmax):
    n = lmax - lmin + 1
    ctx = pkg.Context(0)
    try:
        W = mask_like_W(lmax + 1, seed=20 + kind, cross=(kind == 1))
        ctx.mcm(kind, lmin, lmax, 3.0 * W)
        with _Dev(ctx) as d:
            dW, dM = d.put(W), d.empty(8 * n * n)
            ctx.check(ctx.lib.psx_mcm_d(ctx.handle, kind, lmin, lmax, dW, W.size, dM, n),
                      "psx_mcm_d")
            ref = np.empty((n, n), order="F")
            ctx.d2h(ref, dM)
        got = _nan(n, n)
        ctx.check(ctx.lib.psx_mcm(ctx.handle, kind, lmin, lmax, _ptr(W), W.size, _ptr(got), n),
                  "psx_mcm")
        assert np.array_equal(got, ref), kind
    finally:
        ctx.close()


@pytest.mark.parametrize("lmin,lmax", [(0, 2300), (2, 700)])
def test_pm_host_equals_device(pkg, lmin, lmax):
    n = lmax - lmin + 1
    ctx = pkg.Context(0)
    try:
        W = mask_like_W(lmax + 1, seed=31)
        ctx.mcm_pm(lmin, lmax, 0.5 * W)
        with _Dev(ctx) as d:
            dW, dP, dQ = d.put(W), d.empty(8 * n * n), d.empty(8 * n * n)
            ctx.check(ctx.lib.psx_mcm_pm_d(ctx.handle, lmin, lmax, dW, W.size, dP, dQ, n),
                      "psx_mcm_pm_d")
            rP, rQ = np.empty((n, n), order="F"), np.empty((n, n), order="F")
            ctx.d2h(rP, dP)
            ctx.d2h(rQ, dQ)
        P, Q = _nan(n, n), _nan(n, n)
        ctx.check(ctx.lib.psx_mcm_pm(ctx.handle, lmin, lmax, _ptr(W), W.size, _ptr(P), _ptr(Q),
                                     n), "psx_mcm_pm")
        assert np.array_equal(P, rP) and np.array_equal(Q, rQ)
    finally:
        ctx.close()


def _cov_device(ctx, kind, lmin, lmax, spectra, ratios, W):
    n = lmax - lmin + 1
    with _Dev(ctx) as d:
        sp = (_vp * 4)(*[d.put(s) for s in spectra])
        rp = (_vp * 4)(*([d.put(r) for r in ratios] + [None] * (4 - len(ratios))))
        wp = (_vp * 8)(*([d.put(w) for w in W] + [None] * (8 - len(W))))
        dC = d.empty(8 * n * n)
        ctx.check(ctx.lib.psx_cov_d(ctx.handle, kind, lmin, lmax, sp, rp, wp, W[0].size, 0, dC,
                                    n), "psx_cov_d")
        C = np.empty((n, n), order="F")
        ctx.d2h(C, dC)
    return C


@pytest.mark.parametrize("lmin,lmax", [(2, 2200), (0, 600)])
@pytest.mark.parametrize("kind", [0, 2, 6])
def test_cov_host_equals_device(pkg, kind, lmin, lmax):
    nW, nS, nR = pkg._abi.cov_arity(kind)
    ctx = pkg.Context(0)
    try:
        spectra, ratios = spectra_set(lmax, seed=kind)
        W = [mask_like_W(lmax + 1, seed=40 + i, cross=(i % 2 == 1)) for i in range(nW)]
        ctx.cov(kind, lmin, lmax, spectra[:nS], ratios[:nR], [1.5 * w for w in W])
        ref = _cov_device(ctx, kind, lmin, lmax, spectra[:nS], ratios[:nR], W)
        got = ctx.cov(kind, lmin, lmax, spectra[:nS], ratios[:nR], W)
        assert np.array_equal(got, ref), kind
    finally:
        ctx.close()


@pytest.mark.parametrize("lmax", [2200, 600])
def test_cov_host_domain_error(pkg, lmax):
    """A negative spectrum product under the square root raises through the
    host path, banded (the flag is read after the copy) or not."""
    lmin = 0
    nW, nS, nR = pkg._abi.cov_arity(0)
    ctx = pkg.Context(0)
    try:
        spectra, ratios = spectra_set(lmax, seed=5)
        spectra = [s.copy() for s in spectra]
        bad = (3 * lmax) // 4
        spectra[0][bad] = -spectra[0][bad]
        W = [mask_like_W(lmax + 1, seed=50 + i) for i in range(nW)]
        with pytest.raises(pkg.DomainError):
            ctx.cov(0, lmin, lmax, spectra, ratios, W)
        # the context stays usable
        spectra[0][bad] = -spectra[0][bad]
        C = ctx.cov(0, lmin, lmax, spectra, ratios, W)
        assert np.isfinite(C).all()
    finally:
        ctx.close()


def test_partial_set_host(pkg):
    """TE and PP only: the host path runs the requested blocks one by one."""
    lmax = 2100
    n = lmax + 1
    ctx = pkg.Context(0)
    try:
        Wte, Wp = mask_like_W(n, seed=61, cross=True), mask_like_W(n, seed=62)
        outs = ctx.mcm_set(0, lmax, None, Wte, None, Wp)
        assert outs[0] is None and outs[2] is None
        assert np.array_equal(outs[1], ctx.mcm(1, 0, lmax, Wte))
        P, Q = ctx.mcm_pm(0, lmax, Wp)
        assert np.array_equal(outs[3], P) and np.array_equal(outs[4], Q)
    finally:
        ctx.close()


def test_transfer_buffer_counts_and_yields(pkg):
    """The transfer buffer kept after a host call counts as context memory and
    is given back when a device allocation would otherwise exceed the budget."""
    lmax = 1200
    n = lmax + 1
    W = mask_like_W(n, seed=70)
    probe = pkg.Context(0)
    try:
        base = probe.bytes_in_use()
        probe.mcm(0, 0, lmax, W)
        held = probe.bytes_in_use() - base
        assert held >= 8 * n * n
    finally:
        probe.close()
    # a budget that fits the transfer buffer or a same-size device block, not both
    ctx = pkg.Context(0, hbm_budget=base + held + 8 * n * n // 2)
    try:
        M = ctx.mcm(0, 0, lmax, W)
        p = ctx.alloc(8 * n * n)
        ctx.free(p)
        assert np.array_equal(ctx.mcm(0, 0, lmax, W), M)
    finally:
        ctx.close()


def test_work_buffer_yields_to_allocations(pkg):
    """A transform leaves its device copies in the transfer buffer and its ring
    arrays in the work buffer.  With a budget that holds exactly what the call
    left behind, a device allocation larger than the transfer buffer succeeds
    by taking both back, and the transform then gives the same bits again."""
    from powerspectra_amd.masks import config_masks
    nside, lmax, m = 128, 383, 2
    maps = np.stack(config_masks(nside, 20, (0, 1)))
    probe = pkg.Context(0)
    try:
        base = probe.bytes_in_use()
        probe.mask_spectra(maps, nside, lmax)
        held = probe.bytes_in_use() - base
    finally:
        probe.close()
    npix, nalm = 12 * nside * nside, (lmax + 1) * (lmax + 2) // 2
    xfer = 8 * npix * m + 16 * nalm * m + 8 * m * m * (lmax + 1) + 768
    ctx = pkg.Context(0, hbm_budget=base + held)
    try:
        cl = ctx.mask_spectra(maps, nside, lmax)
        p = ctx.alloc(xfer + 8 * npix)
        ctx.free(p)
        assert np.array_equal(ctx.mask_spectra(maps, nside, lmax), cl)
    finally:
        ctx.close()
