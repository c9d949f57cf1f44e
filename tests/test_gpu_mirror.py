"""GPU: the Julia-API mirror end to end (masks -> SHT -> mcm / coupledcov /
decouple_covmat / `\\`), against the oracle on the same mask spectra."""
import numpy as np
import pytest
from helpers import RTOL_FP64, mask_like_W, norm_err, rel_err, spectra_set

pytestmark = pytest.mark.gpu

NSIDE = 32


def _masks(seed=0):
    """Two smooth cut-sky masks on a HEALPix nside-32 grid (apodised latitude
    cut times a few smoothed holes)."""
    from __graft_entry__ import load_package
    pkg = load_package()
    from powerspectra_amd.sht import RingInfo
    info = RingInfo(NSIDE)
    npix = 12 * NSIDE * NSIDE
    z = np.zeros(npix)
    phi = np.zeros(npix)
    for k in range(info.first.size):
        n = int(info.count[k])
        z[info.first[k]: info.first[k] + n] = info.cth[k]
        phi[info.first[k]: info.first[k] + n] = info.phi0[k] + 2 * np.pi * np.arange(n) / n
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(2):
        b = np.arcsin(z)
        m = 0.5 * (1 + np.tanh((np.abs(b) - 0.3) / 0.1))
        for _ in range(5):
            z0, p0 = rng.uniform(-1, 1), rng.uniform(0, 2 * np.pi)
            cosd = z * z0 + np.sqrt(1 - z * z) * np.sqrt(1 - z0 * z0) * np.cos(phi - p0)
            m *= 1 - np.exp(-(1 - cosd) / 0.005)
        out.append(pkg.HealpixMap(m, NSIDE))
    return out


@pytest.fixture(scope="module")
def masks():
    return _masks(0)


def test_mcm_from_masks_vs_oracle(pkg, oracle, masks):
    m1, m2 = masks
    a1, a2 = pkg.map2alm(m1), pkg.map2alm(m2)
    lmax = a1.lmax
    V = pkg.alm2cl(a1, a2)[: lmax + 1]
    for spec, kind in (("TT", 0), ("TE", 1), ("M⁺⁺", 2), ("M⁻⁻", 3)):
        M = pkg.mcm(spec, a1, a2)
        R = oracle.mcm(kind, 0, lmax, V)
        assert M.axes() == (range(0, lmax + 1), range(0, lmax + 1))
        assert rel_err(M.parent, R) <= RTOL_FP64 and norm_err(M.parent, R) <= RTOL_FP64
    # HealpixMap inputs go through map2alm first (src/modecoupling.jl:250-251)
    M = pkg.mcm("TT", m1, m2, lmin=2)
    R = oracle.mcm(0, 2, lmax, V)
    assert rel_err(M.parent, R) <= RTOL_FP64


def test_block_mcms_and_decoupling(pkg, masks):
    m1, m2 = masks
    a1, a2 = pkg.map2alm(m1), pkg.map2alm(m2)
    lmin = 2
    EE_BB, EB_BE = pkg.mcm(("EE_BB", "EB_BE"), a1, a2, lmin=lmin)
    Mpp = pkg.mcm("M⁺⁺", a1, a2, lmin=lmin).parent
    Mmm = pkg.mcm("M⁻⁻", a1, a2, lmin=lmin).parent
    n = Mpp.shape[0]
    assert np.array_equal(EE_BB.parent, np.block([[Mpp, Mmm], [Mmm, Mpp]]))
    assert np.array_equal(EB_BE.parent, np.block([[Mpp, -Mmm], [-Mmm, Mpp]]))
    # the EE/BB joint solve of maskedalm2spectra (src/modecoupling.jl:362-369)
    rng = np.random.default_rng(0)
    ell = np.arange(lmin, lmin + n)
    clee = pkg.SpectralVector(1e3 / (ell + 1.0) ** 2, lmin - 1)
    clbb = pkg.SpectralVector(1e2 / (ell + 1.0) ** 2, lmin - 1)
    pcl = pkg.block_vector(*(pkg.SpectralVector(v, lmin - 1) for v in (
        Mpp @ clee.parent + Mmm @ clbb.parent, Mmm @ clee.parent + Mpp @ clbb.parent)))
    ee, bb = pkg.spectra(pkg.solve(EE_BB, pcl))
    assert ee.firstindex() == lmin
    assert rel_err(ee.parent, clee.parent) < 1e-9 and rel_err(bb.parent, clbb.parent) < 1e-9
    # single block `\` keeps the offsets (src/blockspectralmatrix.jl:124-129)
    M = pkg.mcm("TT", a1, a2, lmin=lmin)
    x = pkg.solve(M, pkg.SpectralVector(M.parent @ clee.parent, lmin - 1))
    assert x.firstindex() == lmin and rel_err(x.parent, clee.parent) < 1e-9


def test_coupledcov_workspace_vs_oracle(pkg, oracle, masks):
    m1, m2 = masks
    npix = 12 * NSIDE * NSIDE
    rng = np.random.default_rng(3)
    s2a = (rng.uniform(1, 2, npix),) * 3
    s2b = (rng.uniform(1, 2, npix),) * 3
    f1 = pkg.CovField("143_hm1", m1, m2, s2a)
    f2 = pkg.CovField("143_hm2", m2, m1, s2b)
    ws = pkg.CovarianceWorkspace(f1, f2, f1, f2)
    lmax = ws.lmax
    names = ("143_hm1", "143_hm2")
    sp, rt = spectra_set(lmax, 2)
    spectra, ratios = {}, {}
    for XY, base in (("TT", sp[0]), ("EE", sp[1]), ("TE", 0.3 * sp[2])):
        for a in names:
            for b in names:
                spectra[(XY, a, b)] = base * (1.1 if a == b else 1.0)
                ratios[(XY, a, b)] = rt[0] if a == b else np.ones(lmax + 1)
    i, j, p, q = ws.field_names
    W = ws.window_function_W
    E = "∅∅"
    # TTTT through the mirror vs the oracle fed the same spectra, ratios and W
    C = pkg.coupledcov("TT", "TT", ws, spectra, ratios)
    Wk = [W(E, E, i, p, "TT", j, q, "TT"), W(E, E, i, q, "TT", j, p, "TT"),
          W(E, "TT", i, p, "TT", j, q, "TT"), W(E, "TT", j, q, "TT", i, p, "TT"),
          W(E, "TT", i, q, "TT", j, p, "TT"), W(E, "TT", j, p, "TT", i, q, "TT"),
          W("TT", "TT", i, p, "TT", j, q, "TT"), W("TT", "TT", i, q, "TT", j, p, "TT")]
    R = oracle.cov(0, 0, lmax,
                   [spectra["TT", i, p], spectra["TT", j, q], spectra["TT", i, q], spectra["TT", j, p]],
                   [ratios["TT", i, p], ratios["TT", j, q], ratios["TT", i, q], ratios["TT", j, p]], Wk)
    assert rel_err(C.parent, R) <= RTOL_FP64
    # TETE
    C = pkg.coupledcov("TE", "TE", ws, spectra, ratios)
    Wk = [W(E, E, i, p, "TT", j, q, "PP"), W(E, E, i, q, "TP", j, p, "PT"),
          W(E, "PP", i, p, "TT", j, q, "PP"), W(E, "TT", j, q, "PP", i, p, "TT"),
          W("TT", "PP", i, p, "TT", j, q, "PP")]
    R = oracle.cov(3, 0, lmax,
                   [spectra["TT", i, p], spectra["EE", j, q], spectra["TE", i, q], spectra["TE", j, p]],
                   [ratios["TT", i, p], ratios["EE", j, q]], Wk)
    assert rel_err(C.parent, R) <= RTOL_FP64 and norm_err(C.parent, R) <= RTOL_FP64
    # the other pairs run and are symmetric
    for a, b in (("EE", "EE"), ("TT", "TE"), ("TT", "EE"), ("TE", "EE")):
        C = pkg.coupledcov(a, b, ws, spectra, ratios)
        assert np.isfinite(C.parent).all() and np.array_equal(C.parent, C.parent.T)


def test_decouple_covmat_mirror(pkg, masks):
    m1, m2 = masks
    M = pkg.mcm("TT", m1, m2, lmin=2)
    rng = np.random.default_rng(4)
    X = rng.uniform(size=M.shape)
    Cc = pkg.SpectralArray(X.T @ X, firsts=[2, 2])
    Cd = pkg.decouple_covmat(Cc, M, M)
    assert Cd.axes() == Cc.axes()
    back = M.parent @ Cd.parent @ M.parent.T
    assert np.linalg.norm(back - Cc.parent) / np.linalg.norm(Cc.parent) < 1e-9


def test_maskedalm2spectra_block_solves(pkg, ctx, masks):
    """maskedalm2spectra (src/modecoupling.jl:343-380): each spectrum is the
    solve of the right block against the right pseudo-spectrum."""
    from powerspectra_amd.master import maskedalm2spectra
    m1, m2 = masks
    aT1, aT2 = pkg.map2alm(m1), pkg.map2alm(m2)
    aP1, aP2 = pkg.map2alm(m2), pkg.map2alm(m1)           # different T and P masks
    lmax = aT1.lmax
    rng = np.random.default_rng(5)

    def rand_alm():
        a = pkg.Alm(lmax)
        a.alm[:] = rng.standard_normal(a.alm.size) + 1j * rng.standard_normal(a.alm.size)
        return a

    X1 = [rand_alm() for _ in range(3)]
    X2 = [rand_alm() for _ in range(3)]
    lmin = 2
    out = maskedalm2spectra(X1, aT1, aP1, X2, aT2, aP2, lmin=lmin)
    W = lambda a, b: pkg.alm2cl(a, b)[: lmax + 1]
    cl = lambda a, b: pkg.alm2cl(a, b)[lmin: lmax + 1]
    M = lambda k, a, b: ctx.mcm(k, lmin, lmax, W(a, b))
    T, E, B = 0, 1, 2
    for XY, k, (a, b), (x, y) in (("TT", 0, (aT1, aT2), (T, T)), ("TE", 1, (aT1, aP2), (T, E)),
                                  ("ET", 1, (aP1, aT2), (E, T)), ("TB", 1, (aT1, aP2), (T, B)),
                                  ("BT", 1, (aP1, aT2), (B, T))):
        want = np.linalg.solve(M(k, a, b), cl(X1[x], X2[y]))
        got = out[XY]
        assert got.firstindex() == lmin
        assert np.abs(got.parent - want).max() <= 1e-8 * np.abs(want).max()
    Mpp, Mmm = ctx.mcm_pm(lmin, lmax, W(aP1, aP2))
    A = np.block([[Mpp, Mmm], [Mmm, Mpp]])
    want = np.linalg.solve(A, np.concatenate([cl(X1[E], X2[E]), cl(X1[B], X2[B])]))
    n = lmax - lmin + 1
    assert np.abs(out["EE"].parent - want[:n]).max() <= 1e-8 * np.abs(want).max()
    assert np.abs(out["BB"].parent - want[n:]).max() <= 1e-8 * np.abs(want).max()
    A = np.block([[Mpp, -Mmm], [-Mmm, Mpp]])
    want = np.linalg.solve(A, np.concatenate([cl(X1[E], X2[B]), cl(X1[B], X2[E])]))
    assert np.abs(out["EB"].parent - want[:n]).max() <= 1e-8 * np.abs(want).max()
    assert np.abs(out["BE"].parent - want[n:]).max() <= 1e-8 * np.abs(want).max()


def test_master_full_sky_is_identity(pkg):
    """master (src/modecoupling.jl:317-341) on full-sky masks: every mode-coupling
    block is the identity (M-- = 0), so the decoupled spectra are the
    pseudo-spectra of the (I, Q, U) maps."""
    from powerspectra_amd.master import master
    from powerspectra_amd.sht import alm2map_pol, map2alm_pol
    nside = 8
    lmax = 2 * nside          # the iterated transform is exact to ~1e-14 here
    rng = np.random.default_rng(9)
    alms = []
    for _ in range(3):
        a = pkg.Alm(lmax)
        a.alm[:] = rng.standard_normal(a.alm.size) + 1j * rng.standard_normal(a.alm.size)
        for l in range(lmax + 1):
            a[l, 0] = a[l, 0].real
        alms.append(a)
    for a in alms[1:]:
        for m in range(2):
            for l in range(m, 2):
                a[l, m] = 0.0
    maps = [x.pixels for x in alm2map_pol(*alms, nside)]
    ones = np.ones(12 * nside * nside)
    pol = lambda I, Q, U: map2alm_pol(I, Q, U, lmax=lmax, niter=12)
    scal = lambda m: pkg.map2alm(m, lmax=lmax, niter=12)
    out = master(maps, ones, ones, maps, ones, ones, lmin=2, transform=(pol, scal))
    T, E, B = pol(*maps)
    for XY, (a, b) in (("TT", (T, T)), ("TE", (T, E)), ("EE", (E, E)), ("BB", (B, B)),
                       ("EB", (E, B))):
        want = pkg.alm2cl(a, b)[2:]
        assert np.abs(out[XY].parent - want).max() <= 1e-9 * np.abs(want).max()


def test_master_default_path_matches_alm_route(pkg, ctx, masks):
    """master's default path (both masked maps in one polarised analysis, their
    spectra summed on the device, each distinct mask transformed once) against
    the same GPU transforms called one by one through maskedalm2spectra; maps
    masked beforehand (already_masked) give the same spectra bit for bit."""
    from powerspectra_amd.master import maskedalm2spectra, master
    from powerspectra_amd.sht import map2alm, map2alm_pol
    mT, mP = (np.asarray(m.pixels) for m in masks)
    npix = 12 * NSIDE * NSIDE
    rng = np.random.default_rng(21)
    map1 = [rng.standard_normal(npix) for _ in range(3)]
    map2 = [x + 0.3 * rng.standard_normal(npix) for x in map1]
    # the P mask of map 2 is map 1's T mask: one object, transformed once
    out = master(map1, mT, mP, map2, mP, mT, lmin=2, ctx=ctx)
    a1 = map2alm_pol(map1[0] * mT, map1[1] * mP, map1[2] * mP, ctx=ctx)
    a2 = map2alm_pol(map2[0] * mP, map2[1] * mT, map2[2] * mT, ctx=ctx)
    ref = maskedalm2spectra(a1, map2alm(mT, ctx=ctx), map2alm(mP, ctx=ctx), a2,
                            map2alm(mP, ctx=ctx), map2alm(mT, ctx=ctx), lmin=2, ctx=ctx)
    pre = master([map1[0] * mT, map1[1] * mP, map1[2] * mP], mT, mP,
                 [map2[0] * mP, map2[1] * mT, map2[2] * mT], mP, mT,
                 already_masked=True, lmin=2, ctx=ctx)
    assert sorted(out) == sorted(ref) == ["BB", "BE", "BT", "EB", "EE", "ET", "TB", "TE", "TT"]
    for k in out:
        assert out[k].firstindex() == 2 and out[k].lastindex() == 3 * NSIDE - 1
        want = ref[k].parent
        assert np.abs(out[k].parent - want).max() <= 1e-10 * np.abs(want).max(), k
        assert np.array_equal(pre[k].parent, out[k].parent), k


@pytest.mark.parametrize("spec,lmin,lmax", [("TT", 0, 300), ("TE", 2, 257), ("M++", 2, 130),
                                            ("M--", 2, 130)])
def test_mcm_binned_matches_host_product(pkg, ctx, spec, lmin, lmax):
    """mcm_binned (psx_mcm_binned: the block kept on the device, then two
    rocBLAS GEMMs) against P M Q on the host from the unbinned block, with
    binning_matrix's bins (src/util.jl:3-18) and a flat unbinning Q, and
    against P M P^T when Q is left out."""
    V = mask_like_W(lmax + 1, seed=lmax)
    left = np.arange(2, lmax - 20, 30)
    P = pkg.binning_matrix(left, left + 29, lambda l: l * (l + 1.0), lmax)
    Q = (P > 0).T.astype(np.float64)
    M = pkg.mcm_from_spectrum(spec, V, lmin=lmin, lmax=lmax, ctx=ctx).parent
    got = pkg.mcm_binned(spec, V, P, Q, lmin=lmin, lmax=lmax, ctx=ctx)
    want = P[:, lmin:] @ M @ Q[lmin:, :]
    assert got.shape == (P.shape[0], Q.shape[1])
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    got = pkg.mcm_binned(spec, V, P, lmin=lmin, lmax=lmax, ctx=ctx)
    want = P[:, lmin:] @ M @ P[:, lmin:].T
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_mcm_binned_arguments(pkg, ctx):
    """A 2n x 2n block name, and binning matrices of the wrong l length, are refused."""
    V = mask_like_W(101, seed=1)
    P = pkg.binning_matrix(np.arange(2, 90, 10), np.arange(11, 100, 10), lambda l: 1.0, 100)
    with pytest.raises(pkg.ArgumentError):
        pkg.mcm_binned("EE_BB", V, P, lmax=100, ctx=ctx)
    with pytest.raises(pkg.AxisError):
        pkg.mcm_binned("TT", V, P[:, :-1], lmax=100, ctx=ctx)
    with pytest.raises(pkg.AxisError):
        ctx.mcm_binned(0, 0, 100, V, P, P[:, :-1].T)


@pytest.mark.parametrize("lmin,lmax", [(2, 400), (2, 1601), (2, 3071)])
def test_master_decouple_equals_block_path(ctx, lmin, lmax):
    """psx_master_decouple (the five blocks and four concurrent solves on the
    device) against the same blocks through the host path: mcm_set, then
    lu_solve for TT, TE | TB, ET | BT and lu_solve_pm for EE/BB and EB/BE
    (src/modecoupling.jl:348-377).  At lmax 1601 the 2n x 2n system (3200
    rows) takes the multi-workgroup leaves while the n x n ones take the
    single-workgroup leaves, at the same time; lmax 3071 is master's size
    (n = 3070, the 2n system 6140 rows).  (lmin >= 2: below l = 2 the spin-2
    blocks vanish and the EE/BB system is singular, as in Julia.)"""
    from powerspectra_amd._abi import SingularException
    LW = lmax + 1
    Ws = [mask_like_W(LW, 11), mask_like_W(LW, 12, cross=True),
          mask_like_W(LW, 13, cross=True), mask_like_W(LW, 14)]
    n = lmax - lmin + 1
    rng = np.random.default_rng(lmax)
    pcl = rng.standard_normal((9, n)) * np.exp(-np.arange(lmin, lmax + 1) / 300.0)
    got = ctx.master_decouple(lmin, lmax, *Ws, pcl)
    MTT, MTE, MET, Mpp, Mmm = ctx.mcm_set(lmin, lmax, *Ws)
    want = np.empty((9, n))
    want[0] = ctx.lu_solve(MTT, pcl[0])
    want[1:3] = ctx.lu_solve(MTE, pcl[1:3].T).T
    want[3:5] = ctx.lu_solve(MET, pcl[3:5].T).T
    X = ctx.lu_solve_pm(Mpp, Mmm, np.concatenate([pcl[5], pcl[6]]), 1)
    want[5], want[6] = X[:n], X[n:]
    X = ctx.lu_solve_pm(Mpp, Mmm, np.concatenate([pcl[7], pcl[8]]), -1)
    want[7], want[8] = X[:n], X[n:]
    for k in range(9):
        assert np.abs(got[k] - want[k]).max() <= RTOL_FP64 * np.abs(want[k]).max(), k
    # independent check: numpy's LAPACK on the host blocks
    ref = np.linalg.solve(MTT, pcl[0])
    assert np.abs(got[0] - ref).max() <= 1e-9 * np.abs(ref).max()
    A = np.block([[Mpp, -Mmm], [-Mmm, Mpp]])
    ref = np.linalg.solve(A, np.concatenate([pcl[7], pcl[8]]))
    assert np.abs(np.concatenate([got[7], got[8]]) - ref).max() <= 1e-9 * np.abs(ref).max()
    # a second call with all-zero W: every block is singular
    Z = np.zeros(LW)
    with pytest.raises(SingularException):
        ctx.master_decouple(lmin, lmax, Z, Z, Z, Z, pcl)
    again = ctx.master_decouple(lmin, lmax, *Ws, pcl)
    assert np.array_equal(again, got)


def test_master_decouple_arguments(ctx, pkg):
    """psx_master_decouple refuses W spectra of different lengths, pseudo-spectra
    of the wrong shape, lmax below lmin and a null pointer, before any work."""
    import ctypes
    lmax = 40
    W = mask_like_W(lmax + 1, 3)
    pcl = np.ones((9, lmax - 1))
    with pytest.raises(pkg.ArgumentError):
        ctx.master_decouple(2, lmax, W, W, W[:-1], W, pcl)
    with pytest.raises(pkg.AxisError):
        ctx.master_decouple(2, lmax, W, W, W, W, pcl[:8])
    with pytest.raises(pkg.ArgumentError):
        ctx.master_decouple(5, 4, W, W, W, W, np.ones((9, 0)))
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    out = np.empty((9, lmax - 1))
    st = ctx.lib.psx_master_decouple(ctx.handle, 2, lmax, W.size, dp(W), dp(W), dp(W), dp(W),
                                     ctypes.cast(None, ctypes.POINTER(ctypes.c_double)), dp(out))
    assert st == pkg._abi.PSX_ERR_ARGUMENT
