"""GPU: batches of blocks that share one family of 3j products, contracted
against many W spectra in one pass (psx_mcm_batch, psx_cov_batch).

* FP64 batches (vector ALUs, groups of 8) reproduce the single-call kernels;
* the FP64 matrix-core pass agrees with them to FP64 rounding (1e-11
  elementwise where terms cancel, 1e-13 normwise);
* the int8 matrix-core pass (PSX_MIXED: six exact int8 digits of the 3j
  products, seven of the W spectra, 27 digit products, int32 accumulation)
  stays within 1e-9 relative error of the FP64 result on every block,
  elementwise over |ref| >= 1e-6 max|ref|, and within 1e-12 normwise; the
  synthetic cross spectra here flip sign at random multipoles, and where an
  element's terms then cancel (up to 7e4 times in the TE family) its error
  is held to 1e-12 of the magnitude of its terms instead (the FP64 kernels
  themselves carry ~1e-16 of it);
* the mixed pass matches the oracle on sampled rows to 1e-9.
The error of each block is printed."""
import numpy as np
import pytest
from helpers import RTOL_FP64, mask_like_W, norm_err, rel_err, spectra_set

pytestmark = pytest.mark.gpu

RTOL_MIXED = 1e-9
# two FP64 summation orders (matrix-core pass against the vector kernels)
# differ by ~1e-16 of an element's terms; elements whose terms cancel (the
# signed spin-2 families with sign-changing cross spectra) show it at 1e-12
RTOL_ORDER = 1e-11


def _abs_terms(kind, lmin, lmax, W):
    """T[l1, l2] = (2 l2 + 1) kappa sum_l3 |family(l3) What(l3)| for l1 <= l2
    (the magnitudes whose sum is the element, scaled as M), from the closed
    forms: the scale an element's rounding errors are measured against when
    its terms cancel (sign-changing cross spectra, the signed TE family)."""
    n = lmax - lmin + 1
    LW = W.size
    N = 2 * lmax + 3
    g = np.empty(N)
    g[0] = 1.0
    for m in range(1, N):
        g[m] = g[m - 1] * (2 * m - 1) / (2 * m)
    gz = np.concatenate([g, np.zeros(1)])       # index -1 -> 0
    z = (2 * np.arange(N) + 1) * g
    zz = np.concatenate([z, np.zeros(1)])
    Dd = 1.0 / ((2 * np.arange(N) + 1) * g)
    IG = 1.0 / g
    what = np.abs((2 * np.arange(LW) + 1) * W / (4 * np.pi))
    T = np.full((n, n), np.nan)
    for l1 in range(lmin, lmax + 1):
        l2 = np.arange(l1, lmax + 1)[:, None]
        c = np.arange(0, l1 + 1)[None, :]
        S = l1 + l2
        P1, P2 = l1 * (l1 + 1.0), l2 * (l2 + 1.0)
        if kind == 3:
            l3 = S - 2 * c - 1
            p = zz[l1 - c - 1] * z[c] * zz[np.maximum(l2 - c - 1, -1)] * IG[S - c]
            D1 = P1 + P2 - l3 * (l3 + 1.0) - 2.0
            f = p * D1 * D1
        else:
            l3 = S - 2 * c
            p = gz[l1 - c] * g[c] * gz[l2 - c] * Dd[S - c]
            Q = (P1 + P2 - 1.0 - l3 * (l3 + 1.0)) ** 2 - (1.0 + 2.0 * P1 * P2)
            f = p if kind == 0 else p * np.abs(Q) if kind == 1 else p * Q * Q
        w = np.where((l3 >= 0) & (l3 < LW), what[np.clip(l3, 0, LW - 1)], 0.0)
        row = (np.abs(f) * w).sum(axis=1)
        Pi1 = (l1 - 1.0) * l1 * (l1 + 1.0) * (l1 + 2.0)
        Pi2 = (l2[:, 0] - 1.0) * l2[:, 0] * (l2[:, 0] + 1.0) * (l2[:, 0] + 2.0)
        with np.errstate(divide="ignore", invalid="ignore"):
            kap = np.ones_like(Pi2) if kind == 0 else \
                1.0 / (2.0 * np.sqrt(Pi1 * Pi2)) if kind == 1 else 1.0 / (4.0 * Pi1 * Pi2)
        if kind != 0:
            kap = np.where((l1 >= 2) & (l2[:, 0] >= 2), kap, 0.0)
        T[l1 - lmin, l1 - lmin:] = (2.0 * l2[:, 0] + 1.0) * kap * row
    return T


def _mixed_ok(M, R, kind, lmin, lmax, W):
    """Elementwise over |R| >= 1e-6 max|R|: |M - R| <= 1e-9 |R|, or, where the
    element's terms cancel, <= 1e-12 of the magnitude of its terms."""
    e = rel_err(M, R)
    if e <= RTOL_MIXED:
        return e, None
    T = _abs_terms(kind, lmin, lmax, W)
    up = np.triu(np.ones(M.shape, dtype=bool))
    sel = up & (np.abs(R) >= 1e-6 * np.abs(R).max())
    d = np.abs(M - R)[sel]
    ok = (d <= RTOL_MIXED * np.abs(R[sel])) | (d <= 1e-12 * T[sel])
    worst_terms = float((d / T[sel]).max())
    assert ok.all(), (e, worst_terms)
    return e, worst_terms


def _Ws(K, LW, seed):
    """K mask-like spectra: auto spectra, cross spectra with sign changes,
    and overall magnitudes spread over four decades."""
    return [mask_like_W(LW, seed=seed + k, cross=(k % 3 == 1), scale=10.0 ** (-(k % 4)))
            for k in range(K)]


@pytest.mark.parametrize("kind,lmin,lmax,K", [(0, 0, 300, 5), (1, 2, 257, 3), (2, 0, 200, 9),
                                              (3, 0, 180, 2), (0, 0, 96, 16)])
def test_fp64_batch_equals_single_calls(ctx, kind, lmin, lmax, K):
    Ws = _Ws(K, lmax + 1, 3)
    outs = ctx.mcm_batch(kind, lmin, lmax, Ws)
    for W, M in zip(Ws, outs):
        R = ctx.mcm(kind, lmin, lmax, W)
        assert rel_err(M, R) <= 1e-15 and norm_err(M, R) <= 1e-15


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
@pytest.mark.parametrize("lmin,lmax,K,LWx", [(0, 700, 16, 1), (2, 333, 33, 2), (0, 64, 64, 1),
                                             (0, 1, 3, 1), (5, 150, 1, 1)])
def test_matrix_core_batches_against_fp64(pkg, ctx, kind, lmin, lmax, K, LWx):
    """Every block of a PSX_MIXED and a PSX_FP64_MFMA batch against the FP64
    batch; LWx = 2 takes W up to 2 lmax + 1 (all l3 of every family)."""
    abi = pkg._abi
    LW = (2 * lmax + 1) if LWx == 2 else lmax + 1
    Ws = _Ws(K, LW, 10 * kind + K)
    ref = ctx.mcm_batch(kind, lmin, lmax, Ws, abi.FP64)
    mix = ctx.mcm_batch(kind, lmin, lmax, Ws, abi.MIXED)
    f64 = ctx.mcm_batch(kind, lmin, lmax, Ws, abi.FP64_MFMA)
    worst = (0.0, 0.0, 0.0)
    cancel = []
    for k, (R, M, F) in enumerate(zip(ref, mix, f64)):
        e, wt = _mixed_ok(M, R, kind, lmin, lmax, Ws[k])
        if wt is not None:
            cancel.append((k, e, wt))
        nrm, ef = norm_err(M, R), rel_err(F, R)
        worst = tuple(max(a, b) for a, b in zip(worst, (e, nrm, ef)))
        assert nrm <= 1e-12, (k, nrm)
        if wt is None:
            assert ef <= RTOL_ORDER, (k, ef)
        assert norm_err(F, R) <= 1e-13, k
    print(f"kind {kind} lmax {lmax} K {K}: mixed max rel {worst[0]:.2e} normwise {worst[1]:.2e}; "
          f"fp64 mfma max rel {worst[2]:.2e}; blocks with cancelling terms (block, rel, "
          f"err / sum of |terms|): {cancel}", flush=True)


def test_mixed_tt_rows_against_oracle(pkg, ctx, oracle):
    """A 24-spectrum TT batch at lmax 1500 on sampled rows against the oracle
    in long double."""
    abi = pkg._abi
    lmax = 1500
    Ws = _Ws(24, lmax + 1, 77)
    mix = ctx.mcm_batch(0, 0, lmax, Ws, abi.MIXED)
    rows = [(0, 2), (700, 701), (1499, 1500)]
    for W, M in zip(Ws[::5], mix[::5]):
        R = np.full(M.shape, np.nan, order="F")
        for lo, hi in rows:
            oracle.mcm_rows(0, 0, lmax, lo, hi, W, out=R, precision=1)
        # the oracle fills the sampled rows and their mirrored columns
        assert rel_err(M, R) <= RTOL_MIXED


def _cov_args(kind, lmax, nblk, seed):
    nW, nS, nR = {0: (8, 4, 4), 1: (8, 4, 4), 2: (4, 4, 2), 3: (5, 4, 2), 4: (4, 4, 2),
                  5: (4, 4, 2), 6: (2, 4, 0)}[kind]
    sp, rt, ws = [], [], []
    for b in range(nblk):
        s, r = spectra_set(lmax, seed=seed + b)
        sp.append(s[:nS])
        rt.append(r[:nR])
        ws.append([mask_like_W(lmax + 1, seed=100 * seed + 10 * b + w) for w in range(nW)])
    return sp, rt, ws


@pytest.mark.parametrize("kind,nblk,lmax", [(0, 8, 400), (1, 3, 300), (2, 5, 257), (4, 2, 200),
                                            (5, 17, 150), (6, 9, 333), (0, 1, 64)])
def test_cov_batches(pkg, ctx, kind, nblk, lmax):
    """Covariance batches: FP64 equals psx_cov block by block; the matrix-core
    passes against FP64 per block (17 TEEE-Planck blocks take two passes)."""
    abi = pkg._abi
    sp, rt, ws = _cov_args(kind, lmax, nblk, seed=kind + 1)
    ref = ctx.cov_batch(kind, 0, lmax, sp, rt, ws, abi.FP64)
    for b in range(nblk):
        single = ctx.cov(kind, 0, lmax, sp[b], rt[b], ws[b])
        assert rel_err(ref[b], single) <= 1e-15
    mix = ctx.cov_batch(kind, 0, lmax, sp, rt, ws, abi.MIXED)
    f64 = ctx.cov_batch(kind, 0, lmax, sp, rt, ws, abi.FP64_MFMA)
    errs = []
    for b in range(nblk):
        e, nrm = rel_err(mix[b], ref[b]), norm_err(mix[b], ref[b])
        errs.append(e)
        assert e <= RTOL_MIXED and nrm <= RTOL_MIXED, (b, e, nrm)
        assert rel_err(f64[b], ref[b]) <= RTOL_ORDER
        assert np.array_equal(mix[b], mix[b].T)
    print(f"cov kind {kind} lmax {lmax}: mixed per-block max rel "
          + " ".join(f"{e:.1e}" for e in errs), flush=True)


def test_cov_single_call_precisions_and_errors(pkg, ctx):
    """psx_cov takes PSX_MIXED (one block through the matrix-core pass); TETE
    is refused for the matrix cores; a negative spectrum product under a
    square root raises DomainError from the matrix-core epilogue too."""
    abi = pkg._abi
    lmax = 120
    sp, rt, ws = _cov_args(0, lmax, 1, seed=4)
    R = ctx.cov(0, 0, lmax, sp[0], rt[0], ws[0], abi.FP64)
    M = ctx.cov(0, 0, lmax, sp[0], rt[0], ws[0], abi.MIXED)
    assert rel_err(M, R) <= RTOL_MIXED
    sp3, rt3, ws3 = _cov_args(3, lmax, 1, seed=5)
    with pytest.raises(pkg.ArgumentError, match="TETE"):
        ctx.cov(3, 0, lmax, sp3[0], rt3[0], ws3[0], abi.MIXED)
    bad = [s.copy() for s in sp[0]]
    bad[0][50] = -1.0
    with pytest.raises(pkg.DomainError):
        ctx.cov(0, 0, lmax, bad, rt[0], ws[0], abi.MIXED)
    with pytest.raises(pkg.ArgumentError):
        ctx.mcm_batch(0, 0, 10, [mask_like_W(11)] * 65, abi.MIXED)
