"""GPU parity on the workloads' own inputs: seeded cut-sky masks (galactic cut
|b| < 20 deg, holes of 0.5 deg, 1 deg cosine apodisation), mask spectra from
the GPU transform, then the mode-coupling and covariance kernels against the
oracle fed the same spectra."""
import numpy as np
import pytest
from helpers import RTOL_FP64, norm_err, rel_err

pytestmark = pytest.mark.gpu


def _rows_err(A, R, rows):
    m = np.zeros(A.shape, dtype=bool)
    for lo, hi in rows:
        m[lo: hi + 1, :] = True
    a = np.where(m, A, np.nan)
    r = np.where(m, R, np.nan)
    return rel_err(a, r)


def test_tt_lmax767_nside256_full_matrix(pkg, ctx, oracle):
    """Plain TT mode-coupling matrix at lmax 767 from two nside-256 masks with
    50 holes (seeds 0, 1): the whole 768 x 768 matrix."""
    from powerspectra_amd.masks import config_masks, mask_spectra
    lmax = 767
    W = mask_spectra(config_masks(256, 50, (0, 1)), lmax, ctx=ctx)[(0, 1)][: lmax + 1]
    assert W[0] > 0 and np.abs(W[1:]).max() < W[0]
    M = ctx.mcm(0, 0, lmax, W)
    R = oracle.mcm(0, 0, lmax, W)
    assert rel_err(M, R) <= RTOL_FP64 and norm_err(M, R) <= RTOL_FP64


def test_polarised_set_lmax3071_nside1024_rows(pkg, ctx, oracle):
    """TT, TE, ET, ++, -- at lmax 3071 from four nside-1024 masks with 200
    holes (T and P mask per map, seeds 0-3), one fused pass, checked on
    sampled rows of every block against the oracle in long double: with these
    masks the P cross-spectrum is ~1e-6 at odd l against 5 at l = 0, and the
    FP64 recurrence's own error on the -- block reaches 1e-12 at high l while
    the closed form stays within 3e-14 of long double."""
    from powerspectra_amd.masks import config_masks, mask_spectra
    lmax = 3071
    T1, P1, T2, P2 = range(4)
    Wall = mask_spectra(config_masks(1024, 200, (0, 1, 2, 3)), lmax, ctx=ctx)
    W = {k: v[: lmax + 1] for k, v in Wall.items()}
    outs = ctx.mcm_set(0, lmax, W[T1, T2], W[T1, P2], W[P1, T2], W[P1, P2])
    rows = [(0, 3), (700, 701), (1800, 1801), (3069, 3071)]
    for M, (kind, key) in zip(outs, [(0, (T1, T2)), (1, (T1, P2)), (1, (P1, T2)), (2, (P1, P2)),
                                     (3, (P1, P2))]):
        R = np.full(M.shape, np.nan, order="F")
        for lo, hi in rows:
            oracle.mcm_rows(kind, 0, lmax, lo, hi, W[key], out=R, precision=1)
        assert _rows_err(M, R, rows) <= RTOL_FP64, kind


def test_covariance_lmax2508_nside1024_rows(pkg, ctx, oracle):
    """Coupled covariance blocks TTTT, TETE, EEEE, TTEE at lmax 2508 through
    the covariance workspace (effective weights of nside-1024 masks and pixel
    variances, GPU transform) with signal C_l = A (l+10)^-2.5 exp(-(l/2000)^2)
    and white noise N_l = s^2 exp(l(l+1) t^2 / (8 ln 2)); sampled rows."""
    from powerspectra_amd.masks import config_masks
    lmax = 2508
    nside = 1024
    m = config_masks(nside, 200, (0, 1, 2, 3))
    npix = 12 * nside * nside
    rng = np.random.default_rng(0)
    ell = np.arange(lmax + 1, dtype=np.float64)
    theta = np.radians(1.4 / 60.0)
    arcmin = np.radians(1.0 / 60.0)
    fields = []
    for i, name in enumerate(("143_hm1", "143_hm2")):
        sig = rng.uniform(10, 40) * arcmin          # muK rad
        s2 = np.full(npix, sig ** 2 / (4 * np.pi / npix))
        fields.append((pkg.CovField(name, m[2 * i], m[2 * i + 1], (s2, s2, s2)), sig))
    ws = pkg.CovarianceWorkspace(fields[0][0], fields[1][0], fields[0][0], fields[1][0], lmax=lmax)
    A = {"TT": 1e3, "EE": 30.0, "TE": 80.0}
    spectra, ratios = {}, {}
    for XY in ("TT", "EE", "TE"):
        cl = A[XY] * (ell + 10.0) ** -2.5 * np.exp(-(ell / 2000.0) ** 2)
        for a, (fa, sa) in zip(("143_hm1", "143_hm2"), fields):
            for b, (fb, sb) in zip(("143_hm1", "143_hm2"), fields):
                spectra[(XY, a, b)] = cl
                nl = sa * sb * np.exp(ell * (ell + 1) * theta ** 2 / (8 * np.log(2)))
                ratios[(XY, a, b)] = np.sqrt(nl / (sa * sb))
    rows = [(0, 2), (1200, 1201), (2507, 2508)]
    E = "∅∅"
    i, j, p, q = ws.field_names
    W = ws.window_function_W
    cases = {
        ("TT", "TT"): (0, [spectra["TT", i, p], spectra["TT", j, q], spectra["TT", i, q],
                           spectra["TT", j, p]],
                       [ratios["TT", i, p], ratios["TT", j, q], ratios["TT", i, q],
                        ratios["TT", j, p]],
                       [W(E, E, i, p, "TT", j, q, "TT"), W(E, E, i, q, "TT", j, p, "TT"),
                        W(E, "TT", i, p, "TT", j, q, "TT"), W(E, "TT", j, q, "TT", i, p, "TT"),
                        W(E, "TT", i, q, "TT", j, p, "TT"), W(E, "TT", j, p, "TT", i, q, "TT"),
                        W("TT", "TT", i, p, "TT", j, q, "TT"),
                        W("TT", "TT", i, q, "TT", j, p, "TT")]),
        ("TE", "TE"): (3, [spectra["TT", i, p], spectra["EE", j, q], spectra["TE", i, q],
                           spectra["TE", j, p]],
                       [ratios["TT", i, p], ratios["EE", j, q]],
                       [W(E, E, i, p, "TT", j, q, "PP"), W(E, E, i, q, "TP", j, p, "PT"),
                        W(E, "PP", i, p, "TT", j, q, "PP"), W(E, "TT", j, q, "PP", i, p, "TT"),
                        W("TT", "PP", i, p, "TT", j, q, "PP")]),
        ("TT", "EE"): (6, [spectra["TE", i, p], spectra["TE", i, q], spectra["TE", j, q],
                           spectra["TE", j, p]], [],
                       [W(E, E, i, p, "TP", j, q, "TP"), W(E, E, i, q, "TP", j, p, "TP")]),
    }
    for (a, b), (kind, sp, rt, Ws) in cases.items():
        C = pkg.coupledcov(a, b, ws, spectra, ratios).parent
        R = np.full(C.shape, np.nan, order="F")
        for lo, hi in rows:
            R[lo: hi + 1] = oracle.cov(kind, 0, lmax, sp, rt, Ws, rows=(lo, hi))[lo: hi + 1]
        assert _rows_err(C, R, rows) <= RTOL_FP64, (a, b)
    C = pkg.coupledcov("EE", "EE", ws, spectra, ratios).parent
    assert np.isfinite(C).all() and np.array_equal(C, C.T)
