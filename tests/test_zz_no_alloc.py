"""psx_mcm_d_noalloc: the allocation-free mode-coupling entry point, checked
bit for bit against psx_mcm_d through the C ABI on two calls in a row (the
second smaller), with the context's device memory unchanged by the calls."""
import numpy as np
import pytest

from helpers import mask_like_W

pytestmark = pytest.mark.gpu


def _dev_block(ctx, W, n, ld):
    dW = ctx.alloc(W.nbytes)
    dM = ctx.alloc(8 * ld * n)
    ctx.h2d(dW, W)
    return dW, dM


def _fetch(ctx, dM, n, ld):
    full = np.empty((ld, n), order="F")
    ctx.d2h(full, dM)
    return full[:n, :]


def _call(ctx, name, kind, lmin, lmax, dW, LW, dM, ld):
    fn = getattr(ctx.lib, name)
    print(f"{name} kind={kind} lmin={lmin} lmax={lmax} LW={LW} ld={ld}", flush=True)
    ctx.check(fn(ctx.handle, kind, lmin, lmax, dW, LW, dM, ld), name)
    ctx.synchronize()


@pytest.mark.parametrize("calls", [
    # (kind, lmin, lmax, LW, ld): the second call of each pair is smaller
    ((1, 0, 1500, 1501, 1501), (0, 0, 600, 601, 601)),
    ((3, 2, 1100, 900, 1105), (2, 0, 333, 667, 334)),
])
def test_noalloc_matches_mcm_d(ctx, calls):
    for kind, lmin, lmax, LW, ld in calls:
        n = lmax - lmin + 1
        W = mask_like_W(LW, seed=kind + 7 * lmax)
        dW, dM = _dev_block(ctx, W, n, ld)
        dR, dQ = _dev_block(ctx, W, n, ld)
        try:
            # the reference result first, so every cache psx_mcm_d fills exists
            _call(ctx, "psx_mcm_d", kind, lmin, lmax, dR, LW, dQ, ld)
            ref = _fetch(ctx, dQ, n, ld)
            before = ctx.bytes_in_use()
            _call(ctx, "psx_mcm_d_noalloc", kind, lmin, lmax, dW, LW, dM, ld)
            got = _fetch(ctx, dM, n, ld)
            assert ctx.bytes_in_use() == before
            assert np.isfinite(ref).all()
            assert np.array_equal(got, ref), f"kind {kind} lmax {lmax}: not bit-identical"
        finally:
            for p in (dW, dM, dR, dQ):
                ctx.free(p)


def test_noalloc_rejects_beyond_limit(ctx, pkg):
    from importlib import import_module
    abi = import_module(pkg.__name__ + "._abi")
    W = mask_like_W(64)
    dW = ctx.alloc(W.nbytes)
    try:
        ctx.h2d(dW, W)
        fn = ctx.lib.psx_mcm_d_noalloc
        # lmax above the static scratch's limit: refused on the host, before
        # any device work (the output pointer is never touched)
        st = fn(ctx.handle, 0, 0, 8192, dW, W.size, dW, 8193)
        assert st == abi.PSX_ERR_ARGUMENT
        assert "8191" in ctx.lib.psx_ctx_last_error(ctx.handle).decode()
        st = fn(ctx.handle, 9, 0, 10, dW, W.size, dW, 11)
        assert st == abi.PSX_ERR_ARGUMENT
    finally:
        ctx.free(dW)
