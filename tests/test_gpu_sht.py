"""GPU: spherical-harmonic transforms of HEALPix maps (psx_map2alm /
psx_alm2map) against the host transforms of the same definition, and
round trips at nside 1024 where the Legendre start values underflow double
range and the scaled recurrence matters."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _smooth_map(nside, seed):
    from powerspectra_amd.sht import RingInfo
    info = RingInfo(nside)
    npix = 12 * nside * nside
    z = np.zeros(npix)
    phi = np.zeros(npix)
    for k in range(info.first.size):
        n = int(info.count[k])
        z[info.first[k]: info.first[k] + n] = info.cth[k]
        phi[info.first[k]: info.first[k] + n] = info.phi0[k] + 2 * np.pi * np.arange(n) / n
    rng = np.random.default_rng(seed)
    m = 0.5 * (1 + np.tanh((np.abs(np.arcsin(z)) - 0.3) / 0.1))
    for _ in range(4):
        z0, p0 = rng.uniform(-1, 1), rng.uniform(0, 2 * np.pi)
        cosd = z * z0 + np.sqrt(1 - z * z) * np.sqrt(1 - z0 * z0) * np.cos(phi - p0)
        m *= 1 - np.exp(-(1 - cosd) / 0.02)
    return m + 0.1 * rng.standard_normal(npix)


def _rand_alm(pkg, lmax, seed):
    rng = np.random.default_rng(seed)
    a = pkg.Alm(lmax)
    a.alm[:] = rng.standard_normal(a.alm.size) + 1j * rng.standard_normal(a.alm.size)
    for l in range(lmax + 1):
        a[l, 0] = a[l, 0].real
    return a


@pytest.mark.parametrize("nside,lfac,niter", [(1, 2, 0), (2, 3, 1), (4, 2, 3), (8, 3, 0),
                                              (8, 4, 3), (16, 2, 3), (32, 3, 3)])
def test_map2alm_vs_host(pkg, ctx, nside, lfac, niter):
    from powerspectra_amd.sht import map2alm_host
    lmax = lfac * nside - (1 if lfac == 3 else 0)
    m = _smooth_map(nside, nside + lfac)
    got = ctx.map2alm(m, nside, lmax, niter)[0]
    want = map2alm_host(m, lmax=lmax, niter=niter).alm
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("nside,lmax", [(1, 2), (4, 11), (16, 40), (32, 95)])
def test_alm2map_vs_host(pkg, ctx, nside, lmax):
    from powerspectra_amd.sht import alm2map_host
    a = _rand_alm(pkg, lmax, nside)
    got = ctx.alm2map(a.alm, nside, lmax)[0]
    want = alm2map_host(a, nside).pixels
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("nside,lmax,niter,nmaps", [(16, 47, 3, 3), (16, 47, 3, 6),
                                                     (1024, 3071, 1, 5)])
def test_batch_of_maps(pkg, ctx, nside, lmax, niter, nmaps):
    """Maps go through the Legendre passes in groups of up to four that share
    the recurrence (6 = 4 + 2, 5 = 4 + 1); each result equals the one-map
    transform bit for bit, in both directions (nside 1024: two ring-pair
    chunks per m)."""
    maps = np.stack([_smooth_map(nside, s) for s in range(nmaps)])
    A = ctx.map2alm(maps, nside, lmax, niter)
    for i in range(nmaps):
        assert np.array_equal(A[i], ctx.map2alm(maps[i], nside, lmax, niter)[0]), i
    B = ctx.alm2map(A, nside, lmax)
    for i in range(nmaps):
        assert np.array_equal(B[i], ctx.alm2map(A[i], nside, lmax)[0]), i


def test_roundtrip_nside1024(pkg, ctx):
    """Band-limited alms (lmax = 2 nside) survive synthesis + analysis: with
    ten refinements every coefficient converges to ~1e-11 (the slowest are
    m = 0, as for any HEALPix quadrature)."""
    nside, lmax = 1024, 2048
    a = _rand_alm(pkg, lmax, 7)
    t0 = time.perf_counter()
    m = ctx.alm2map(a.alm, nside, lmax)[0]
    t1 = time.perf_counter()
    b = ctx.map2alm(m, nside, lmax, 10)[0]
    t2 = time.perf_counter()
    print(f"nside {nside} lmax {lmax}: alm2map {t1 - t0:.3f} s, map2alm(niter=10) {t2 - t1:.3f} s")
    assert np.abs(b - a.alm).max() <= 1e-10 * np.abs(a.alm).max()


def test_scaled_legendre_start_vs_host(pkg, ctx):
    """nside 64, lmax 191: near the poles lambda_mm ~ sin^m theta starts below
    2^-600 for the highest m and is carried scaled; the result equals the
    host transform (which evaluates the same recurrence unscaled)."""
    from powerspectra_amd.sht import alm2map_host, map2alm_host
    nside, lmax = 64, 191
    a = _rand_alm(pkg, lmax, 3)
    m = ctx.alm2map(a.alm, nside, lmax)[0]
    mh = alm2map_host(a, nside).pixels
    assert np.abs(m - mh).max() <= 1e-12 * np.abs(mh).max()
    b = ctx.map2alm(mh, nside, lmax, 1)[0]
    bh = map2alm_host(mh, lmax=lmax, niter=1).alm
    assert np.abs(b - bh).max() <= 1e-12 * np.abs(bh).max()


def test_mcm_from_masks_uses_gpu_transform(pkg, ctx, oracle):
    """mcm(:TT, mask1, mask2) end to end: GPU map2alm, alm2cl, GPU MCM, against
    the oracle fed the host transform's mask spectrum."""
    from powerspectra_amd.sht import map2alm_host
    nside = 32
    m1, m2 = _smooth_map(nside, 1), _smooth_map(nside, 2)
    M = pkg.mcm("TT", m1, m2)
    a1, a2 = map2alm_host(m1), map2alm_host(m2)
    lmax = a1.lmax
    V = pkg.alm2cl(a1, a2)[: lmax + 1]
    R = oracle.mcm(0, 0, lmax, V)
    from helpers import rel_err
    assert rel_err(M.parent, R) <= 1e-10


def test_nside2048_roundtrip_and_lmax6143_masks(pkg, ctx):
    """nside 2048 (polar rings up to 8188 pixels): a band-limited round trip
    at lmax 4096, and the lmax-6143 workload's mask spectra (two nside-2048
    masks with 500 holes) finite, with W_0 = 4 pi fsky^2-like and the steep
    fall-off of an apodised mask."""
    from powerspectra_amd.masks import config_masks, mask_spectra
    nside, lmax = 2048, 4096
    a = _rand_alm(pkg, lmax, 11)
    m = ctx.alm2map(a.alm, nside, lmax)[0]
    b = ctx.map2alm(m, nside, lmax, 10)[0]
    assert np.abs(b - a.alm).max() <= 1e-9 * np.abs(a.alm).max()
    t0 = time.perf_counter()
    W = mask_spectra(config_masks(nside, 500, (0, 1)), 6143, ctx=ctx)[(0, 1)]
    print(f"nside 2048 masks -> W (lmax 6143): {time.perf_counter() - t0:.2f} s")
    assert W.size == 6144 and np.isfinite(W).all()
    assert 0.5 * 4 * np.pi * 0.6 ** 2 < W[0] < 4 * np.pi
    assert np.abs(W[3000:]).max() < 1e-4 * W[0]


# ---- spin 2 -------------------------------------------------------------------

def _rand_pol_alms(pkg, lmax, seed):
    """T, E, B with real m = 0 parts and E = B = 0 for l < 2."""
    T, E, B = (_rand_alm(pkg, lmax, seed + i) for i in range(3))
    for a in (E, B):
        for m in range(min(2, lmax + 1)):
            for l in range(m, min(2, lmax + 1)):
                a[l, m] = 0.0
    return T, E, B


@pytest.mark.parametrize("nside,lfac,niter", [(1, 2, 0), (2, 3, 1), (4, 2, 3), (8, 3, 0),
                                              (8, 4, 2), (16, 2, 3), (32, 3, 1)])
def test_map2alm_pol_vs_host(pkg, ctx, nside, lfac, niter):
    """Polarised analysis (T, E, B) against the numpy transform of the same
    definition (Wigner-d recurrences, pinned against sympy and the analytic
    quadrupole on the CPU)."""
    from powerspectra_amd.sht import map2alm_pol_host
    lmax = lfac * nside - (1 if lfac == 3 else 0)
    I, Q, U = (_smooth_map(nside, 10 * nside + lfac + i) for i in range(3))
    got = ctx.map2alm_pol(np.stack([I, Q, U]), nside, lmax, niter)[0]
    want = [a.alm for a in map2alm_pol_host(I, Q, U, lmax=lmax, niter=niter)]
    for g, w in zip(got, want):
        assert np.abs(g - w).max() <= 1e-12 * np.abs(w).max()


def test_map2alm_pol_several_maps_equal_single_calls(pkg, ctx):
    """Three (I, Q, U) maps in one call (the I maps copied first, every Q and U
    copied in while the I maps are transformed) equal one call per map bit for
    bit."""
    nside, lmax, niter = 16, 47, 2
    maps = np.stack([np.stack([_smooth_map(nside, 700 + 3 * k + i) for i in range(3)])
                     for k in range(3)])
    both = ctx.map2alm_pol(maps, nside, lmax, niter)
    for k in range(3):
        one = ctx.map2alm_pol(maps[k], nside, lmax, niter)[0]
        assert np.array_equal(both[k], one), k


@pytest.mark.parametrize("nside,lmax", [(1, 2), (4, 11), (16, 40), (32, 95)])
def test_alm2map_pol_vs_host(pkg, ctx, nside, lmax):
    from powerspectra_amd.sht import alm2map_pol_host
    T, E, B = _rand_pol_alms(pkg, lmax, nside)
    got = ctx.alm2map_pol(np.stack([T.alm, E.alm, B.alm]), nside, lmax)[0]
    want = [m.pixels for m in alm2map_pol_host(T, E, B, nside)]
    for g, w in zip(got, want):
        assert np.abs(g - w).max() <= 1e-12 * np.abs(w).max()


def test_pol_quadrupole_analytic(pkg, ctx):
    """a_E(2, 0) = 1 gives Q = -sqrt(15/(32 pi)) sin^2(theta), U = 0; a_B(2, 0)
    = 1 gives U = -sqrt(15/(32 pi)) sin^2(theta), Q = 0 (HEALPix convention)."""
    from powerspectra_amd.sht import RingInfo
    nside, lmax = 16, 8
    info = RingInfo(nside)
    s2 = np.empty(12 * nside * nside)
    for k in range(info.first.size):
        s2[info.first[k]: info.first[k] + info.count[k]] = 1.0 - info.cth[k] ** 2
    want = -np.sqrt(15 / (32 * np.pi)) * s2
    E = pkg.Alm(lmax)
    E[2, 0] = 1.0
    Z = pkg.Alm(lmax)
    _, Q, U = ctx.alm2map_pol(np.stack([Z.alm, E.alm, Z.alm]), nside, lmax)[0]
    assert np.abs(Q - want).max() <= 1e-14 and np.abs(U).max() <= 1e-14
    _, Q, U = ctx.alm2map_pol(np.stack([Z.alm, Z.alm, E.alm]), nside, lmax)[0]
    assert np.abs(U - want).max() <= 1e-14 and np.abs(Q).max() <= 1e-14


def test_pol_scaled_start_vs_host(pkg, ctx):
    """nside 64, lmax 191: the Wigner-d start values of the highest m fall
    below 2^-600 near the poles and are carried scaled."""
    from powerspectra_amd.sht import alm2map_pol_host, map2alm_pol_host
    nside, lmax = 64, 191
    T, E, B = _rand_pol_alms(pkg, lmax, 21)
    got = ctx.alm2map_pol(np.stack([T.alm, E.alm, B.alm]), nside, lmax)[0]
    want = [m.pixels for m in alm2map_pol_host(T, E, B, nside)]
    for g, w in zip(got[1:], want[1:]):
        assert np.abs(g - w).max() <= 1e-12 * np.abs(w).max()
    a = ctx.map2alm_pol(np.stack(want), nside, lmax, 1)[0]
    h = map2alm_pol_host(*want, lmax=lmax, niter=1)
    for g, w in zip(a[1:], h[1:]):
        assert np.abs(g - w.alm).max() <= 1e-12 * np.abs(w.alm).max()


def test_pol_roundtrip_nside1024(pkg, ctx):
    """Band-limited (T, E, B) at lmax = 2 nside survive polarised synthesis and
    analysis with ten refinements; a pure-E input leaves B at the same level."""
    nside, lmax = 1024, 2048
    T, E, B = _rand_pol_alms(pkg, lmax, 31)
    t0 = time.perf_counter()
    maps = ctx.alm2map_pol(np.stack([T.alm, E.alm, B.alm]), nside, lmax)[0]
    t1 = time.perf_counter()
    teb = ctx.map2alm_pol(maps, nside, lmax, 10)[0]
    t2 = time.perf_counter()
    print(f"pol nside {nside} lmax {lmax}: alm2map_pol {t1 - t0:.3f} s, "
          f"map2alm_pol(niter=10) {t2 - t1:.3f} s")
    for g, w in zip(teb, (T, E, B)):
        assert np.abs(g - w.alm).max() <= 1e-10 * np.abs(w.alm).max()
    Z = pkg.Alm(lmax)
    maps = ctx.alm2map_pol(np.stack([Z.alm, E.alm, Z.alm]), nside, lmax)[0]
    teb = ctx.map2alm_pol(maps, nside, lmax, 10)[0]
    assert np.abs(teb[2]).max() <= 1e-10 * np.abs(E.alm).max()


# ---- mask spectra on the device ----------------------------------------------

@pytest.mark.parametrize("nside,lmax,nmaps", [(1, 2, 2), (16, 47, 1), (32, 95, 3), (8, 20, 8)])
def test_mask_spectra_vs_host_alm2cl(pkg, ctx, nside, lmax, nmaps):
    """psx_mask_spectra (map2alm + all cross-spectra on the device) against the
    host alm2cl of the same GPU alms."""
    maps = np.stack([_smooth_map(nside, 40 + s) for s in range(nmaps)])
    C = ctx.mask_spectra(maps, nside, lmax, 2)
    A = ctx.map2alm(maps, nside, lmax, 2)
    alms = [pkg.Alm(lmax, lmax, a) for a in A]
    for i in range(nmaps):
        for j in range(nmaps):
            want = pkg.alm2cl(alms[i], alms[j])
            assert np.abs(C[i, j] - want).max() <= 1e-13 * np.abs(want).max(), (i, j)


def test_mask_spectra_config_masks_and_limits(pkg, ctx):
    """Two nside-256 config masks at lmax 767 through masks.mask_spectra, and
    more than 8 maps per device call refused."""
    from powerspectra_amd.masks import config_masks, mask_spectra
    m = config_masks(256, 50, (0, 1))
    W = mask_spectra(m, 767, ctx=ctx)
    A = ctx.map2alm(np.stack(m), 256, 767, 3)
    a0, a1 = pkg.Alm(767, 767, A[0]), pkg.Alm(767, 767, A[1])
    want = pkg.alm2cl(a0, a1)
    assert np.abs(W[(0, 1)] - want).max() <= 1e-13 * np.abs(want).max()
    assert np.array_equal(W[(0, 1)], W[(1, 0)])
    with pytest.raises(pkg.ArgumentError):
        ctx.mask_spectra(np.zeros((9, 12 * 4 * 4)), 4, 8, 0)


@pytest.mark.parametrize("lmax,nsets", [(0, 1), (5, 2), (100, 6), (767, 8), (2900, 2)])
def test_alm2cl_device_vs_host(pkg, ctx, lmax, nsets):
    """psx_alm2cl (all cross-spectra of coefficient sets anywhere in host
    memory) against the host alm2cl of the same sets; a set passed twice gives
    its auto-spectrum, bit for bit, in all four entries.  At lmax 2900 each
    set is 64 MB, longer than a staging buffer and not a whole number of the
    512 KB lines it is recut into."""
    rng = np.random.default_rng(lmax + nsets)
    nalm = (lmax + 1) * (lmax + 2) // 2
    alms = [pkg.Alm(lmax, lmax, rng.standard_normal(nalm) + 1j * rng.standard_normal(nalm))
            for _ in range(nsets)]
    C = ctx.alm2cl([a.alm for a in alms], lmax)
    assert C.shape == (nsets, nsets, lmax + 1)
    for i in range(nsets):
        for j in range(nsets):
            want = pkg.alm2cl(alms[i], alms[j])
            assert np.abs(C[i, j] - want).max() <= 1e-13 * np.abs(want).max(), (i, j)
    D = ctx.alm2cl([alms[0].alm, alms[0].alm], lmax)
    for i in range(2):
        for j in range(2):
            assert np.array_equal(D[i, j], C[0, 0])


@pytest.mark.parametrize("nside,lmax,repeat", [(8, 20, False), (16, 47, True)])
def test_master_spectra_vs_separate_calls(pkg, ctx, nside, lmax, repeat):
    """psx_master_spectra (masking, both analyses and the spectra on the
    device) against the same steps as separate calls: host products,
    psx_map2alm_pol, psx_alm2cl and psx_mask_spectra.  A mask address given
    twice (repeat) is transformed once; masked maps with already_masked give
    the same result bit for bit."""
    rng = np.random.default_rng(nside)
    maps = [rng.standard_normal(12 * nside * nside) for _ in range(6)]
    masks = [_smooth_map(nside, 60 + s) for s in range(4)]
    if repeat:
        masks[3] = masks[0]
    pcl, W = ctx.master_spectra(maps, masks, nside, lmax, 2)
    which = (0, 1, 1, 2, 3, 3)
    masked = [m * masks[which[i]] for i, m in enumerate(maps)]
    teb = ctx.map2alm_pol(np.stack(masked).reshape(2, 3, -1), nside, lmax, 2)
    P = ctx.alm2cl([teb[0, 0], teb[0, 1], teb[0, 2], teb[1, 0], teb[1, 1], teb[1, 2]], lmax)
    C = ctx.mask_spectra(np.stack(masks), nside, lmax, 2)
    for x in range(3):
        for y in range(3):
            want = P[x, 3 + y]
            assert np.abs(pcl[x, y] - want).max() <= 1e-13 * np.abs(want).max(), (x, y)
    for a in range(2):
        for b in range(2):
            want = C[a, 2 + b]
            assert np.abs(W[a, b] - want).max() <= 1e-13 * np.abs(want).max(), (a, b)
    pcl2, W2 = ctx.master_spectra(masked, masks, nside, lmax, 2, already_masked=True)
    assert np.array_equal(pcl2, pcl) and np.array_equal(W2, W)
    with pytest.raises(pkg.ArgumentError):
        ctx.master_spectra(maps[:5], masks, nside, lmax)
    with pytest.raises(pkg.ArgumentError):
        ctx.master_spectra(maps, masks, nside, 4 * nside + 1)


def test_alm2cl_device_limits(pkg, ctx):
    """More than 8 sets, no set, a length that does not match lmax and a
    negative lmax are refused."""
    a = np.zeros(6, dtype=np.complex128)
    with pytest.raises(pkg.ArgumentError):
        ctx.alm2cl([a] * 9, 2)
    with pytest.raises(pkg.ArgumentError):
        ctx.alm2cl([], 2)
    with pytest.raises(pkg.ArgumentError):
        ctx.alm2cl([a[:5]], 2)
    with pytest.raises(pkg.ArgumentError):
        ctx.alm2cl([a[:0]], -1)
