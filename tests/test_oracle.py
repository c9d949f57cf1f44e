"""CPU tests that pin the oracle (oracle/libpsx_oracle.so) before it is
trusted as the parity reference for the HIP path.

What pins what:
  * Wigner families  -- exact sympy values (tests/golden/wigner_golden.json,
    written by tests/golden/make_wigner_golden.py), the sum rule, the sign at
    l3 = l1+l2, (l l 0; m -m 0)^2 = 1/(2l+1), double vs long double.
  * Closed forms the GPU kernels are built on -- checked against sympy at
    small l and against the oracle's recurrence at large l (mpmath tables).
  * Mode-coupling loops -- brute force from the sympy values; full-sky and
    flat-W closed forms; the reference's own NaMaster diagonal fixtures
    (test/data/mcm_{TT,EE,TE}_diag.txt, copied to tests/golden/reference_data):
    W fitted from the TT diagonal must reproduce the TT diagonal and predict
    the EE and TE diagonals.
  * Covariance loops -- full-sky Knox form; brute force from sympy values.
  * decouple_covmat -- the reference KAT (test/test_covmat.jl:11-17).
"""
import json
import os

import mpmath
import numpy as np
import pytest
from helpers import mask_like_W, rel_err, spectra_set

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


@pytest.fixture(scope="module")
def golden():
    with open(os.path.join(GOLD, "wigner_golden.json")) as f:
        return json.load(f)["families"]


def _gold_family(golden, l1, l2, m1, m2):
    for fam in golden:
        if (fam["l1"], fam["l2"], fam["m1"], fam["m2"]) == (l1, l2, m1, m2):
            return {int(k): mpmath.mpf(v) for k, v in fam["values"].items()}
    return None


# ---------------------------------------------------------------- families --
def test_family_vs_sympy(oracle, golden):
    worst = 0.0
    for fam in golden:
        l1, l2, m1, m2 = fam["l1"], fam["l2"], fam["m1"], fam["m2"]
        nmin, f = oracle.wigner_family(l1, l2, m1, m2)
        scale = max(abs(float(v)) for v in fam["values"].values()) or 1.0
        for k, v in fam["values"].items():
            err = abs(f[int(k) - nmin] - float(v)) / scale
            worst = max(worst, err)
            assert err < 1e-13, (l1, l2, m1, m2, k, f[int(k) - nmin], v)
    assert worst < 1e-13


@pytest.mark.parametrize("l1,l2,m1,m2", [(1000, 1500, -2, 2), (3000, 3071, -2, 2), (2500, 2500, -2, 2),
                                         (5, 3000, -2, 2), (2, 6000, -2, 2), (3071, 3071, 0, 0),
                                         (100, 4000, 0, 0), (6143, 6143, -2, 2)])
def test_family_invariants_large_l(oracle, l1, l2, m1, m2):
    nmin, f = oracle.wigner_family(l1, l2, m1, m2)
    _, hi, lo = oracle.wigner_family(l1, l2, m1, m2, "ld")
    j = np.arange(nmin, nmin + f.size)
    assert abs(np.sum((2 * j + 1) * f * f) - 1.0) < 1e-13
    assert np.max(np.abs(f - (hi + lo))) / np.max(np.abs(hi)) < 1e-13
    # sign of the last entry, (l3 = l1 + l2): (-1)^(l1 - l2 - m1 - m2)
    assert np.sign(f[-1]) == (-1) ** ((l1 - l2 - m1 - m2) % 2)


def test_family_special_values(oracle):
    for l in (0, 1, 2, 7, 100, 1999):
        for m in (0, 2):
            if m > l:
                continue
            nmin, f = oracle.wigner_family(l, l, -m, m)
            assert nmin == 0
            assert abs(f[0] ** 2 - 1.0 / (2 * l + 1)) < 1e-15
    # 00 family vanishes for odd l1 + l2 + l3
    nmin, f = oracle.wigner_family(30, 47, 0, 0)
    j = np.arange(nmin, nmin + f.size)
    assert np.all(f[(30 + 47 + j) % 2 == 1] == 0.0)
    # invalid |m| > l: zeros (the physical value; unpinned in the reference)
    nmin, f = oracle.wigner_family(1, 5, -2, 2)
    assert np.all(f == 0.0)


# ------------------------------------------- closed forms (GPU algorithm) --
def _gammas(n):
    mpmath.mp.dps = 40
    g = [mpmath.mpf(1)]
    for m in range(1, n):
        g.append(g[-1] * (2 * m - 1) / (2 * m))
    return g


def _closed(l1, l2, l3, g):
    """(l3 l1 l2; 0 0 0)^2, (l3 l1 l2;0 0 0)(l3 l1 l2;0 -2 2), (l3 l1 l2; 0 -2 2)^2."""
    J = l1 + l2 + l3
    P1, P2 = l1 * (l1 + 1), l2 * (l2 + 1)
    Pi = mpmath.mpf(P1 * P2 * (l1 + 2) * (l1 - 1) * (l2 + 2) * (l2 - 1))
    if J % 2 == 0:
        G = J // 2
        f00sq = g[G - l1] * g[G - l2] * g[G - l3] / ((2 * G + 1) * g[G])
        D0 = P1 + P2 - l3 * (l3 + 1)
        Q = D0 * (D0 - 2) - 2 * P1 * P2
        return f00sq, f00sq * Q / (2 * mpmath.sqrt(Pi)), f00sq * Q * Q / (4 * Pi)
    G = (J - 1) // 2
    z = lambda m: (2 * m + 1) * g[m]
    D1 = P1 + P2 - l3 * (l3 + 1) - 2
    return mpmath.mpf(0), mpmath.mpf(0), D1 * D1 * z(G - l1) * z(G - l2) * z(G - l3) / (4 * Pi * g[G + 1])


def test_closed_forms_vs_sympy(golden):
    g = _gammas(64)
    for l1 in range(2, 13):
        for l2 in range(l1, 13):
            a = _gold_family(golden, l1, l2, 0, 0)
            b = _gold_family(golden, l1, l2, -2, 2)
            for l3 in range(abs(l1 - l2), l1 + l2 + 1):
                c00, c02, c22 = _closed(l1, l2, l3, g)
                assert abs(c00 - a[l3] ** 2) < 1e-24
                assert abs(c02 - a[l3] * b[l3]) < 1e-24
                assert abs(c22 - b[l3] ** 2) < 1e-24


@pytest.mark.parametrize("l1,l2", [(700, 1100), (1500, 1500), (37, 2900), (2, 3071)])
def test_closed_forms_vs_recurrence_large_l(oracle, l1, l2):
    g = _gammas(l1 + l2 + 3)
    nmin, f0 = oracle.wigner_family(l1, l2, 0, 0)
    _, f2 = oracle.wigner_family(l1, l2, -2, 2)
    rng = np.random.default_rng(l1)
    for l3 in sorted(set(rng.integers(nmin, l1 + l2 + 1, 40).tolist() + [nmin, l1 + l2])):
        c00, c02, c22 = _closed(l1, l2, l3, g)
        k = l3 - nmin
        scale = max(np.abs(f2).max() ** 2, 1e-300)
        assert abs(float(c00) - f0[k] ** 2) <= 1e-13 * np.abs(f0).max() ** 2
        assert abs(float(c02) - f0[k] * f2[k]) <= 1e-13 * np.abs(f0).max() * np.abs(f2).max()
        assert abs(float(c22) - f2[k] ** 2) <= 1e-13 * scale


# ------------------------------------------------------------ MCM loops --
def _bruteforce_mcm(golden, kind, lmax, W):
    mpmath.mp.dps = 30
    n = lmax + 1
    M = np.zeros((n, n))
    for l1 in range(n):
        for l2 in range(l1, n):
            a = _gold_family(golden, min(l1, l2), max(l1, l2), 0, 0)
            b = (_gold_family(golden, min(l1, l2), max(l1, l2), -2, 2)
                 if min(l1, l2) >= 2 else None)
            xi = mpmath.mpf(0)
            for l3 in range(abs(l1 - l2), min(l1 + l2, W.size - 1) + 1):
                even = (l1 + l2 + l3) % 2 == 0
                if kind == 0:
                    w = a[l3] ** 2
                elif kind == 1:
                    w = a[l3] * b[l3] if (b is not None and even) else 0
                elif kind == 2:
                    w = b[l3] ** 2 if (b is not None and even) else 0
                else:
                    w = b[l3] ** 2 if (b is not None and not even) else 0
                xi += (2 * l3 + 1) * w * mpmath.mpf(float(W[l3]))
            xi /= 4 * mpmath.pi
            M[l1, l2] = float((2 * l2 + 1) * xi)
            M[l2, l1] = float((2 * l1 + 1) * xi)
    return M


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
@pytest.mark.parametrize("LW", [13, 25])
def test_mcm_bruteforce_from_sympy(oracle, golden, kind, LW):
    lmax = 12
    W = mask_like_W(LW, seed=LW + kind, cross=True)
    R = _bruteforce_mcm(golden, kind, lmax, W)
    M = oracle.mcm(kind, 0, lmax, W)
    assert rel_err(M, R) < 1e-13
    M2 = oracle.mcm(kind, 3, lmax, W)
    assert np.array_equal(M2, M[3:, 3:])


def test_mcm_full_sky_and_flat(oracle):
    lmax = 150
    W = np.zeros(lmax + 1)
    W[0] = 4 * np.pi
    n = lmax + 1
    I2 = np.diag((np.arange(n) >= 2).astype(float))
    assert np.abs(oracle.mcm(0, 0, lmax, W) - np.eye(n)).max() < 1e-13
    assert np.abs(oracle.mcm(1, 0, lmax, W) - I2).max() < 1e-13
    assert np.abs(oracle.mcm(2, 0, lmax, W) - I2).max() < 1e-13
    assert np.abs(oracle.mcm(3, 0, lmax, W)).max() < 1e-13
    c = 0.71
    Wf = np.full(2 * lmax + 1, 4 * np.pi * c)
    target = c * (2 * np.arange(n) + 1.0)[None, :] * np.ones((n, 1))
    assert rel_err(oracle.mcm(0, 0, lmax, Wf), target) < 1e-13
    pm = oracle.mcm(2, 0, lmax, Wf) + oracle.mcm(3, 0, lmax, Wf)
    assert rel_err(pm[2:, 2:], target[2:, 2:]) < 1e-13
    # (2l2+1)/(2l1+1) symmetry
    M = oracle.mcm(2, 0, lmax, mask_like_W(lmax + 1, seed=4))
    ell = 2.0 * np.arange(n) + 1
    X = M / ell[None, :]
    assert rel_err(X, X.T) < 1e-15


def _namaster_fit():
    """Least-squares W (even l3 <= 767) from the reference's TT diagonal."""
    from oracle import oracle as O
    D = os.path.join(GOLD, "reference_data")
    tt = np.loadtxt(os.path.join(D, "mcm_TT_diag.txt"))
    ee = np.loadtxt(os.path.join(D, "mcm_EE_diag.txt"))
    te = np.loadtxt(os.path.join(D, "mcm_TE_diag.txt"))
    lmax = 767
    ells = np.arange(2, 767)

    def design(kind):
        A = np.zeros((ells.size, lmax + 1))
        for i, l in enumerate(ells):
            nmin, f0 = O.wigner_family(l, l, 0, 0)
            _, f2 = O.wigner_family(l, l, -2, 2)
            w = {"TT": f0 * f0, "EE": f2 * f2, "TE": f0 * f2}[kind]
            l3 = np.arange(nmin, nmin + w.size)
            sel = (l3 <= lmax) & (l3 % 2 == 0)
            A[i, l3[sel]] = (2 * l + 1) / (4 * np.pi) * (2 * l3[sel] + 1) * w[sel]
        return A

    cols = np.arange(0, lmax + 1, 2)
    Att = design("TT")[:, cols]
    Wev, *_ = np.linalg.lstsq(Att, tt, rcond=None)
    W = np.zeros(lmax + 1)
    W[cols] = Wev
    return W, tt, ee, te, design


def test_reference_namaster_diagonals(oracle):
    """The reference checks mcm(:TT/:M⁺⁺/:TE, mask, mask) diagonals against
    NaMaster (test/test_mcm.jl:12-50) at rtol sqrt(eps); its mask is not in the
    repository.  All three diagonals depend only on the even-l3 mask spectrum,
    so: fit it from the 765 TT values (384 unknowns), then the oracle's MCM
    diagonals must reproduce TT and predict EE (M⁺⁺) and TE."""
    W, tt, ee, te, _ = _namaster_fit()
    lmax = 767
    d = {}
    for kind, name in ((0, "TT"), (2, "EE"), (1, "TE")):
        M = oracle.mcm(kind, 0, lmax, W)
        d[name] = np.diag(M)[2:767]
    assert np.max(np.abs(d["TT"] - tt) / np.abs(tt)) < 1e-12
    assert np.max(np.abs(d["EE"] - ee) / np.abs(ee)) < 5e-7
    assert np.max(np.abs(d["TE"] - te) / np.abs(te)) < 5e-7
    # a plausible mistake (the TE sign, or using the odd parity for EE) fails
    assert np.max(np.abs(-d["TE"] - te) / np.abs(te)) > 1.0
    mm = np.diag(oracle.mcm(3, 0, lmax, W))[2:767]
    assert np.max(np.abs(mm - ee) / np.abs(ee)) > 0.5


# ----------------------------------------------------------- covariance --
def test_cov_full_sky_knox(oracle):
    lmax = 120
    spec, rat = spectra_set(lmax, 1)
    w = np.array([1.0, 0.9, 0.3, 0.25, 0.2, 0.15, 0.05, 0.04])
    W = []
    for k in range(8):
        x = np.zeros(lmax + 1)
        x[0] = 4 * np.pi * w[k]
        W.append(x)
    C = oracle.cov(0, 0, lmax, spec, rat, W)
    ip, jq, iq, jp = spec
    rip, rjq, riq, rjp = rat
    l = np.arange(lmax + 1)
    diag = (np.sqrt(ip * ip * jq * jq) * w[0] + np.sqrt(iq * iq * jp * jp) * w[1]
            + np.sqrt(ip * ip) * w[2] * rjq * rjq + np.sqrt(jq * jq) * w[3] * rip * rip
            + np.sqrt(iq * iq) * w[4] * rjp * rjp + np.sqrt(jp * jp) * w[5] * riq * riq
            + w[6] * rip * rjq * rip * rjq + w[7] * riq * rjp * riq * rjp) / (2 * l + 1)
    assert rel_err(np.diag(C), diag) < 1e-14
    assert np.abs(C - np.diag(np.diag(C))).max() == 0.0


@pytest.mark.parametrize("kind", range(7))
def test_cov_bruteforce_from_sympy(oracle, golden, kind):
    """Each covariance loop rebuilt from exact sympy families and the
    reference's coefficient algebra (src/covariance.jl:110-441)."""
    lmax = 10
    nW = {0: 8, 1: 8, 2: 4, 3: 5, 4: 4, 5: 4, 6: 2}[kind]
    nR = {0: 4, 1: 4, 2: 2, 3: 2, 4: 2, 5: 2, 6: 0}[kind]
    spec, rat = spectra_set(lmax, kind)
    rat = rat[:nR]
    W = [mask_like_W(lmax + 1, seed=10 * kind + k) for k in range(nW)]
    C = oracle.cov(kind, 0, lmax, spec, rat, W)
    mpmath.mp.dps = 30

    def xi(Wk, w, l1, l2, parity):
        s = mpmath.mpf(0)
        for l3 in range(abs(l1 - l2), min(l1 + l2, lmax) + 1):
            ev = (l1 + l2 + l3) % 2 == 0
            if parity == "even" and not ev:
                continue
            s += (2 * l3 + 1) * w(l3) * mpmath.mpf(float(Wk[l3]))
        return float(s / (4 * mpmath.pi))

    for l1 in range(2, lmax + 1):
        for l2 in range(l1, lmax + 1):
            a = _gold_family(golden, l1, l2, 0, 0)
            b = _gold_family(golden, l1, l2, -2, 2)
            f00 = lambda l3: a[l3] ** 2
            f22 = lambda l3: b[l3] ** 2
            f02 = lambda l3: a[l3] * b[l3]
            S, R = spec, rat
            x, y = l1, l2
            sq = lambda v: np.sqrt(v)
            if kind in (0, 1):
                fam, par = (f00, "all") if kind == 0 else (f22, "even")
                X = [xi(W[k], fam, x, y, par) for k in range(8)]
                ip, jq, iq, jp = S
                rip, rjq, riq, rjp = R
                ref = (sq(ip[x] * ip[y] * jq[x] * jq[y]) * X[0] + sq(iq[x] * iq[y] * jp[x] * jp[y]) * X[1]
                       + sq(ip[x] * ip[y]) * X[2] * rjq[x] * rjq[y] + sq(jq[x] * jq[y]) * X[3] * rip[x] * rip[y]
                       + sq(iq[x] * iq[y]) * X[4] * rjp[x] * rjp[y] + sq(jp[x] * jp[y]) * X[5] * riq[x] * riq[y]
                       + X[6] * rip[x] * rjq[x] * rip[y] * rjq[y] + X[7] * riq[x] * rjp[x] * riq[y] * rjp[y])
            elif kind == 2:
                X = [xi(W[k], f00, x, y, "all") for k in range(4)]
                TTip, TTjp, TEiq, TEjq = S
                rip, rjp = R
                ref = (sq(TTip[x] * TTip[y]) * (TEjq[x] + TEjq[y]) * X[0]
                       + sq(TTjp[x] * TTjp[y]) * (TEiq[x] + TEiq[y]) * X[1]
                       + (TEjq[x] + TEjq[y]) * X[2] * rip[x] * rip[y]
                       + (TEiq[x] + TEiq[y]) * X[3] * rjp[x] * rjp[y]) / 2
            elif kind == 3:
                TTip, EEjq, TEiq, TEjp = S
                rTT, rPP = R
                X1 = xi(W[0], f02, x, y, "even")
                X2 = xi(W[1], f00, x, y, "all")
                X3 = xi(W[2], f02, x, y, "even")
                X4 = xi(W[3], f02, x, y, "even")
                X5 = xi(W[4], f02, x, y, "even")
                ref = (sq(TTip[x] * TTip[y] * EEjq[x] * EEjq[y]) * X1
                       + 0.5 * (TEiq[x] * TEjp[y] + TEjp[x] * TEiq[y]) * X2
                       + sq(TTip[x] * TTip[y]) * X3 * rPP[x] * rPP[y]
                       + sq(EEjq[x] * EEjq[y]) * X4 * rTT[x] * rTT[y]
                       + X5 * rTT[x] * rTT[y] * rPP[x] * rPP[y])
            elif kind in (4, 5):
                fam = f02 if kind == 4 else f22
                X = [xi(W[k], fam, x, y, "even") for k in range(4)]
                EEjq, EEjp, TEip, TEiq = S
                rjq, rjp = R
                ref = (sq(EEjq[x] * EEjq[y]) * (TEip[x] + TEip[y]) * X[0]
                       + sq(EEjp[x] * EEjp[y]) * (TEiq[x] + TEiq[y]) * X[1]
                       + (TEip[x] + TEip[y]) * X[2] * rjq[x] * rjq[y]
                       + (TEiq[x] + TEiq[y]) * X[3] * rjp[x] * rjp[y]) / 2
            else:
                X = [xi(W[k], f00, x, y, "all") for k in range(2)]
                TEip, TEiq, TEjq, TEjp = S
                ref = ((TEip[x] * TEjq[y] + TEjq[x] * TEip[y]) * X[0]
                       + (TEiq[x] * TEjp[y] + TEjp[x] * TEiq[y]) * X[1]) / 2
            assert abs(C[x, y] - ref) <= 1e-13 * abs(ref) + 1e-300, (kind, x, y, C[x, y], ref)
            assert C[y, x] == C[x, y]


# --------------------------------------------------------------- linalg --
def test_decouple_reference_kat(oracle):
    """test/test_covmat.jl:11-17."""
    A = np.array([[1.0, 0.2, 0.3], [0.4, 2.0, 0.15], [0.3, 0.1, 1.44]])
    B1 = np.array([[1.2, 0.6, 0.1], [0.3, 1.4, 0.5], [0.44, 0.2, 1.3]])
    B2 = np.array([[1.6, 0.4, 0.1], [0.3, 1.4, 0.9], [0.45, 0.8, 1.7]])
    Cref = np.linalg.inv(B1) @ A @ np.linalg.inv(B2).T
    C = oracle.decouple_covmat(A, B1, B2)
    assert np.allclose(C, Cref, rtol=1.5e-8, atol=0)


def test_lu_solve_vs_numpy(oracle):
    rng = np.random.default_rng(1)
    A = rng.standard_normal((60, 60)) + 5 * np.eye(60)
    B = rng.standard_normal((60, 2))
    X = oracle.lu_solve(A, B)
    assert np.allclose(X, np.linalg.solve(A, B), rtol=1e-12, atol=1e-12)
    with pytest.raises(RuntimeError):
        oracle.lu_solve(np.zeros((3, 3)), np.ones(3))


# ---------------------------------------------------------------- QuickPol --
def _qp_cells(lmax, bl, bu):
    for lpp in range(2, lmax + 1):
        for l in range(max(2, lpp - bl), min(lmax, lpp + bu) + 1):
            yield lpp, l


def test_quickpol_oracle_tt_reduction(oracle):
    """s = nu = 0: both families are (l' l l''; 0 0 0), so Ξ[l'', l] equals the
    (independently pinned) TT mode-coupling sum M[l, l''] / (2 l'' + 1) with
    V = 4 pi W / (2 l' + 1)."""
    lmax, bl, bu = 40, 5, 7
    rng = np.random.default_rng(3)
    W = rng.uniform(0.1, 1.0, 2 * lmax + 1)
    band = oracle.quickpol_xi(0, 0, 0, 0, lmax, bl, bu, W)
    ell = np.arange(W.size)
    M = oracle.mcm(0, 0, lmax, 4 * np.pi * W / (2 * ell + 1))
    for lpp, l in _qp_cells(lmax, bl, bu):
        want = M[l, lpp] / (2 * lpp + 1)
        assert abs(band[bu + lpp - l, l] - want) <= 1e-14 * abs(want) + 1e-16


@pytest.mark.parametrize("nu1,nu2,s1,s2", [(1, -2, 2, 0), (0, 3, -2, 2), (-4, 2, 2, -2),
                                            (2, 2, 0, 2)])
def test_quickpol_oracle_bruteforce_sympy(oracle, nu1, nu2, s1, s2):
    """Every written cell against exact sympy 3j sums, the sign
    (-1)^(s1+s2+nu1+nu2), and entries in rows/columns 0, 1 (not written by the
    loop) scaled by that sign only (src/beam.jl:98-99)."""
    from sympy.physics.wigner import wigner_3j
    lmax, bl, bu = 7, 2, 3
    rng = np.random.default_rng(nu1 + 10 * s1 + 100)
    W = rng.uniform(-1.0, 1.0, 2 * lmax + 1)
    start = np.asfortranarray(rng.uniform(1.0, 2.0, (bl + bu + 1, lmax + 1)))
    band = oracle.quickpol_xi(nu1, nu2, s1, s2, lmax, bl, bu, W, band=start.copy(order="F"))
    sgn = (-1) ** ((s1 + s2 + nu1 + nu2) % 2)
    for lpp in range(0, lmax + 1):
        for l in range(max(0, lpp - bl), min(lmax, lpp + bu) + 1):
            got = band[bu + lpp - l, l]
            if lpp < 2 or l < 2:
                assert got == sgn * start[bu + lpp - l, l]
                continue
            want = 0.0
            for lp in range(abs(l - lpp), l + lpp + 1):
                a = wigner_3j(lp, l, lpp, s1 + nu1, -s1, -nu1)
                b = wigner_3j(lp, l, lpp, s2 + nu2, -s2, -nu2)
                if a != 0 and b != 0:
                    want += W[lp] * float(a * b)
            assert abs(got - sgn * want) <= 1e-14, (lpp, l, got, sgn * want)
