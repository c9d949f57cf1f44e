"""GPU: the persistent request loop (psx_server).  Results of a mix of
mode-coupling and covariance requests of different lmax are bit-identical to
each request run alone and to the single-call entry points; admission follows
the arena; failures stay with their own request."""
import numpy as np
import pytest
from helpers import RTOL_FP64, mask_like_W, rel_err, spectra_set

pytestmark = pytest.mark.gpu


def _requests(pkg, srv, lmax_list, seed=0):
    """Put spectra for a mix of requests and return their submit arguments."""
    abi = pkg._abi
    rng = np.random.default_rng(seed)
    reqs = []
    key = 1000 * (seed + 1)
    for i, lmax in enumerate(lmax_list):
        LW = lmax + 1
        t = i % 5
        if t == 0:    # TT
            key += 1
            srv.put(key, mask_like_W(LW, seed=i))
            reqs.append(dict(rtype=abi.REQ_MCM, kind=abi.MCM_00, lmin=0, lmax=lmax, W=[key]))
        elif t == 1:  # TE
            key += 1
            srv.put(key, mask_like_W(LW, seed=i, cross=True))
            reqs.append(dict(rtype=abi.REQ_MCM, kind=abi.MCM_02, lmin=2, lmax=lmax, W=[key]))
        elif t == 2:  # ++ and --
            key += 1
            srv.put(key, mask_like_W(LW, seed=i))
            reqs.append(dict(rtype=abi.REQ_MCM_PM, lmin=0, lmax=lmax, W=[key]))
        elif t == 3:  # the full set
            ks = []
            for c in (False, True, True, False):
                key += 1
                srv.put(key, mask_like_W(LW, seed=10 * i + len(ks), cross=c))
                ks.append(key)
            reqs.append(dict(rtype=abi.REQ_MCM_SET, lmin=0, lmax=lmax, W=ks))
        else:         # covariance TTTT
            sp, rt = spectra_set(lmax, seed=i)
            ws, ss, rs = [], [], []
            for k in range(8):
                key += 1
                srv.put(key, mask_like_W(LW, seed=100 * i + k))
                ws.append(key)
            for v in sp:
                key += 1
                srv.put(key, v)
                ss.append(key)
            for v in rt:
                key += 1
                srv.put(key, v)
                rs.append(key)
            reqs.append(dict(rtype=abi.REQ_COV, kind=abi.COV_TTTT, lmin=0, lmax=lmax, W=ws,
                             spectra=ss, ratios=rs))
    return reqs


def _outs(pkg, r):
    n = r["lmax"] - r["lmin"] + 1
    return [np.empty((n, n), order="F") for _ in range(pkg._abi.Server._NOUT[r["rtype"]])]


@pytest.fixture(scope="module")
def server(pkg, ctx):
    s = pkg.Server(ctx, slots=4, arena_bytes=1 << 30, lmax_max=512)
    yield s
    s.close()


def test_mix_equals_alone_and_single_calls(pkg, ctx, server, oracle):
    lmaxs = [300, 97, 211, 150, 128, 64, 257, 180, 33, 120]
    reqs = _requests(pkg, server, lmaxs)
    alone = []
    for r in reqs:
        out = _outs(pkg, r)
        t = server.submit(**r, out=out)
        st, times = server.wait(t)
        assert st == 0 and times[2] >= times[1] >= 0
        server.release(t)
        alone.append(out)
    # the same requests all at once (several in flight on the device)
    mixed = [_outs(pkg, r) for r in reqs]
    tickets = [server.submit(**r, out=o) for r, o in zip(reqs, mixed)]
    for t in tickets:
        assert server.wait(t)[0] == 0
        server.release(t)
    for a, m in zip(alone, mixed):
        for x, y in zip(a, m):
            assert np.array_equal(x, y)
    # identical to the single-call entry points, and to the oracle
    r = reqs[0]
    lmax = r["lmax"]
    W = mask_like_W(lmax + 1, seed=0)
    M = ctx.mcm(0, 0, lmax, W)
    assert np.array_equal(M, alone[0][0])
    assert rel_err(alone[0][0], oracle.mcm(0, 0, lmax, W)) <= RTOL_FP64
    r = reqs[2]
    Wp = mask_like_W(r["lmax"] + 1, seed=2)
    P, Q = ctx.mcm_pm(0, r["lmax"], Wp)
    assert np.array_equal(P, alone[2][0]) and np.array_equal(Q, alone[2][1])
    r = reqs[4]
    sp, rt = spectra_set(r["lmax"], seed=4)
    Ws = [mask_like_W(r["lmax"] + 1, seed=400 + k) for k in range(8)]
    C = ctx.cov(0, 0, r["lmax"], sp, rt, Ws)
    assert np.array_equal(C, alone[4][0])
    assert rel_err(C, oracle.cov(0, 0, r["lmax"], sp, rt, Ws)) <= RTOL_FP64


def test_device_held_results_fetch_release(pkg, server):
    reqs = _requests(pkg, server, [90, 140, 75], seed=1)
    ref = []
    for r in reqs:
        out = _outs(pkg, r)
        t = server.submit(**r, out=out)
        server.wait(t)
        server.release(t)
        ref.append(out)
    tickets = [server.submit(**r) for r in reqs]
    for t, want in zip(tickets, ref):
        assert server.wait(t)[0] == 0
        got = server.fetch(t)
        server.release(t)
        for x, y in zip(got, want):
            assert np.array_equal(x, y)
    with pytest.raises(pkg.ArgumentError):
        server.fetch(tickets[0])          # released


def test_admission_through_a_small_arena(pkg, ctx):
    abi = pkg._abi
    lmax = 200
    n = lmax + 1
    # room for one 201^2 block plus its scratch, not two
    s = pkg.Server(ctx, slots=4, arena_bytes=8 * n * n + (1 << 20), lmax_max=256)
    try:
        W = mask_like_W(n, seed=5)
        s.put(7, W)
        outs = [np.empty((n, n), order="F") for _ in range(6)]
        ts = [s.submit(abi.REQ_MCM, abi.MCM_00, 0, lmax, W=[7], out=[o]) for o in outs]
        for t in ts:
            assert s.wait(t)[0] == 0
            s.release(t)
        for o in outs[1:]:
            assert np.array_equal(o, outs[0])
        # a request larger than the arena is refused at submit
        with pytest.raises(pkg.BudgetError):
            s.submit(abi.REQ_MCM_SET, 0, 0, lmax, W=[7, 7, 7, 7])
        with pytest.raises(pkg.ArgumentError):
            s.submit(abi.REQ_MCM, abi.MCM_00, 0, 300, W=[7])      # above lmax_max
        # a device-held result keeps its arena space: the next request waits in
        # the queue (holding its spectra, which cannot be evicted) until the
        # first is released
        s.put(8, W)
        a = s.submit(abi.REQ_MCM, abi.MCM_00, 0, lmax, W=[7])
        assert s.wait(a)[0] == 0
        b = s.submit(abi.REQ_MCM, abi.MCM_00, 0, lmax, W=[8], out=[outs[1]])
        assert not s.poll(b)
        with pytest.raises(pkg.ArgumentError):
            s.evict(8)
        assert np.array_equal(s.fetch(a)[0], outs[0])
        s.release(a)
        assert s.wait(b)[0] == 0
        s.release(b)
        s.evict(8)
    finally:
        s.close()


def test_errors_stay_with_their_request(pkg, server):
    abi = pkg._abi
    lmax = 80
    n = lmax + 1
    server.put(1, mask_like_W(n, seed=1))
    with pytest.raises(pkg.ArgumentError):
        server.submit(abi.REQ_MCM, abi.MCM_00, 0, lmax, W=[424242])   # not resident
    with pytest.raises(pkg.ArgumentError):
        server.submit(abi.REQ_MCM, 9, 0, lmax, W=[1])                   # unknown kind
    # a covariance request with a negative spectrum product fails with
    # DomainError, the requests around it do not
    sp, rt = spectra_set(lmax, seed=3)
    sp[0] = sp[0].copy()
    sp[0][5] = -sp[0][5]      # TT_ip[5] * TT_ip[l] < 0 for every l != 5
    keys = []
    for k, v in enumerate([mask_like_W(n, seed=k) for k in range(8)] + list(sp) + list(rt)):
        server.put(500 + k, v)
        keys.append(500 + k)
    good1 = server.submit(abi.REQ_MCM, abi.MCM_00, 0, lmax, W=[1])
    bad = server.submit(abi.REQ_COV, abi.COV_TTTT, 0, lmax, W=keys[:8], spectra=keys[8:12],
                        ratios=keys[12:16])
    good2 = server.submit(abi.REQ_MCM, abi.MCM_00, 0, lmax, W=[1])
    assert server.wait(good1)[0] == 0
    st, _ = server.wait(bad, raise_on_error=False)
    assert st == abi.PSX_ERR_DOMAIN
    with pytest.raises(pkg.DomainError):
        server.wait(bad)
    assert server.wait(good2)[0] == 0
    for t in (good1, bad, good2):
        server.release(t)
    server.evict(500)


def test_context_close_refused_while_server_alive(pkg):
    c = pkg.Context(0)
    s = pkg.Server(c, slots=1, arena_bytes=1 << 24, lmax_max=64)
    try:
        with pytest.raises(pkg.ArgumentError, match="live server"):
            c.close()
        # the context is still usable
        W = mask_like_W(65, seed=3)
        c.mcm(0, 0, 64, W)
    finally:
        s.close()
    c.close()
    assert c.handle is None
