"""GPU parity: FP64 LU (the library's blocked getrf and substitution, rocBLAS
updates) and decoupling through the C ABI.

Reference: decouple_covmat (src/covariance.jl:8-14) and `\\`
(src/blockspectralmatrix.jl:101-129).
"""
import numpy as np
import pytest
from helpers import mask_like_W, rel_err

pytestmark = pytest.mark.gpu


def test_decouple_reference_kat(ctx):
    """test/test_covmat.jl:11-17 of the reference, matrices inline there."""
    A = np.array([[1.0, 0.2, 0.3], [0.4, 2.0, 0.15], [0.3, 0.1, 1.44]])
    B1 = np.array([[1.2, 0.6, 0.1], [0.3, 1.4, 0.5], [0.44, 0.2, 1.3]])
    B2 = np.array([[1.6, 0.4, 0.1], [0.3, 1.4, 0.9], [0.45, 0.8, 1.7]])
    Cref = np.linalg.inv(B1) @ A @ np.linalg.inv(B2).T
    C = ctx.decouple_covmat(A, B1, B2)
    assert np.allclose(C, Cref, rtol=1.5e-8, atol=0)
    assert rel_err(C, Cref) < 1e-13


def test_decouple_matches_oracle_and_numpy(ctx, oracle):
    rng = np.random.default_rng(2)
    n = 400
    Y = rng.standard_normal((n, n))
    Y = Y @ Y.T
    B1 = np.eye(n) + 0.1 * rng.standard_normal((n, n))
    B2 = np.eye(n) + 0.1 * rng.standard_normal((n, n))
    C = ctx.decouple_covmat(Y, B1, B2)
    R = oracle.decouple_covmat(Y, B1, B2)
    N = np.linalg.solve(B1, Y) @ np.linalg.inv(B2).T
    assert np.linalg.norm(C - R) / np.linalg.norm(R) < 1e-11
    assert np.linalg.norm(C - N) / np.linalg.norm(N) < 1e-11


def test_decouple_invariant_with_gpu_mcm(ctx):
    """test/test_covmat.jl:115-127: C_coupled = M C_dec M^T for a TT MCM M."""
    lmax = 300
    M = ctx.mcm(0, 0, lmax, mask_like_W(lmax + 1, seed=9))
    rng = np.random.default_rng(0)
    X = rng.uniform(size=M.shape)
    Cc = X.T @ X
    Cd = ctx.decouple_covmat(Cc, M, M)
    back = M @ Cd @ M.T
    assert np.linalg.norm(back - Cc) / np.linalg.norm(Cc) < 1e-9


def test_lu_solve(ctx, oracle):
    rng = np.random.default_rng(5)
    n = 513
    A = np.eye(n) * 3 + rng.standard_normal((n, n))
    b = rng.standard_normal(n)
    x = ctx.lu_solve(A, b)
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 1e-12
    B = rng.standard_normal((n, 3))
    X = ctx.lu_solve(A, B)
    R = oracle.lu_solve(A, B)
    assert np.linalg.norm(X - R) / np.linalg.norm(R) < 1e-11


def test_singular(ctx, pkg):
    A = np.ones((4, 4))
    with pytest.raises(pkg.SingularException):
        ctx.lu_solve(A, np.ones(4))
    with pytest.raises(pkg.SingularException):
        ctx.decouple_covmat(np.eye(4), np.zeros((4, 4)), np.eye(4))
    with pytest.raises(pkg.AxisError):
        ctx.lu_solve(np.eye(4), np.ones(3))


@pytest.mark.parametrize("n", [1, 5, 16, 17, 32, 33, 256, 257, 289, 300, 513, 1000, 1025, 2100,
                               3072, 3100, 4100])
def test_lu_factor_vs_lapack(ctx, n):
    """The blocked getrf (leaf kernels + rocBLAS updates) against LAPACK's
    getrf (scipy): the same row interchanges and factors to rounding.  Up to
    3072 leaf rows one workgroup factors a 16-column leaf in registers; above
    that each column is a launch over many workgroups (n = 3100 and 4100 run
    both, switching when the remaining rows drop to 3072)."""
    import scipy.linalg as sl
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    LU, ipiv = ctx.lu_factor(A)
    lu_ref, piv_ref = sl.lu_factor(A)
    assert np.array_equal(ipiv - 1, piv_ref)
    assert np.linalg.norm(LU - lu_ref) / np.linalg.norm(lu_ref) < 1e-12


def test_lu_factor_mcm_and_solve(ctx):
    """A TT mode-coupling matrix at lmax 1500 (no row interchanges expected
    beyond LAPACK's): factors as LAPACK, and `M \\ b` through psx_lu_solve."""
    import scipy.linalg as sl
    lmax = 1500
    M = ctx.mcm(0, 0, lmax, mask_like_W(lmax + 1, seed=3))
    LU, ipiv = ctx.lu_factor(M)
    lu_ref, piv_ref = sl.lu_factor(M)
    assert np.array_equal(ipiv - 1, piv_ref)
    assert np.linalg.norm(LU - lu_ref) / np.linalg.norm(lu_ref) < 1e-12
    b = M @ (1.0 / (np.arange(lmax + 1) + 1.0) ** 2)
    x = ctx.lu_solve(M, b)
    assert np.linalg.norm(x - sl.lu_solve((lu_ref, piv_ref), b)) / np.linalg.norm(x) < 1e-12


def test_lu_factor_singular(ctx, pkg):
    with pytest.raises(pkg.SingularException):
        ctx.lu_factor(np.ones((40, 40)))


@pytest.mark.parametrize("n,k", [(2100, 2000), (3100, 5), (4100, 3500), (4100, 700)])
def test_lu_factor_zero_column_position(ctx, pkg, n, k):
    """An exactly zero column k stays zero through every update, so the first
    zero pivot is column k + 1 (LAPACK's info): found by the one-workgroup
    leaf (n = 2100), by the per-column leaf (n = 3100, k = 5), and by each of
    them after the other has run (n = 4100).  The solve refuses it too."""
    rng = np.random.default_rng(n + k)
    A = rng.standard_normal((n, n))
    A[:, k] = 0.0
    with pytest.raises(pkg.SingularException, match=rf"position {k + 1} "):
        ctx.lu_factor(A)
    with pytest.raises(pkg.SingularException, match=rf"position {k + 1} "):
        ctx.lu_solve(A, np.ones(n))


def test_lu_factor_structured_pivots(ctx):
    """Pivoting that moves rows across leaves and blocks: a matrix whose
    largest entry of every column sits in the bottom rows (LAPACK swaps it up
    at every step), against scipy."""
    import scipy.linalg as sl
    n = 600
    rng = np.random.default_rng(1)
    A = rng.uniform(0.0, 1.0, (n, n))
    A[n - 64:, :] += 10.0 * np.arange(1, 65)[:, None]
    LU, ipiv = ctx.lu_factor(A)
    lu_ref, piv_ref = sl.lu_factor(A)
    assert np.array_equal(ipiv - 1, piv_ref)
    assert np.linalg.norm(LU - lu_ref) / np.linalg.norm(lu_ref) < 1e-12


def test_lu_solve_device_buffers(ctx):
    """psx_lu_solve_d on device buffers equals psx_lu_solve on host buffers."""
    rng = np.random.default_rng(4)
    n = 700
    A = np.asfortranarray(np.eye(n) * 2 + rng.standard_normal((n, n)))
    b = rng.standard_normal(n)
    want = ctx.lu_solve(A, b)
    dA, dB = ctx.alloc(8 * n * n), ctx.alloc(8 * n)
    try:
        ctx.h2d(dA, A.ravel(order="F"))
        ctx.h2d(dB, b)
        ctx.check(ctx.lib.psx_lu_solve_d(ctx.handle, n, dA, n, dB, n, 1), "lu_solve_d")
        got = np.empty(n)
        ctx.d2h(got, dB)
    finally:
        ctx.free(dA)
        ctx.free(dB)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 255, 256, 257, 320, 511, 777, 2100])
def test_lu_solve_substitution_sizes(ctx, n):
    """Solves with 1-4 right-hand sides run the library's forward and back
    substitution (128-row diagonal blocks, one launch per block); 5 run its
    many-right-hand-side triangular solve (64-row diagonal blocks, rocBLAS
    updates).  Both against LAPACK (scipy lu_solve)."""
    import scipy.linalg as sl
    rng = np.random.default_rng(100 + n)
    A = rng.standard_normal((n, n)) + np.diag(rng.uniform(1.0, 2.0, n)) * np.sqrt(n)
    ref_lu = sl.lu_factor(A)
    for nrhs in (1, 2, 4, 5):
        B = rng.standard_normal((n, nrhs))
        X = ctx.lu_solve(A, B)
        R = sl.lu_solve(ref_lu, B)
        assert np.linalg.norm(X - R) / np.linalg.norm(R) < 1e-12, (n, nrhs)


def test_lu_solve_substitution_padded_lda(ctx):
    """psx_lu_solve_d with lda and ldb larger than n, and pivoting across
    blocks.  The matrix has cond ~ 6e6 (numpy and scipy solves differ by
    5e-12), so the check is the normwise backward error
    |A x - b| / (|A| |x| + |b|), which a stable LU solve keeps near machine
    epsilon whatever the conditioning, and the forward error against scipy
    within cond x 1e-15."""
    import scipy.linalg as sl
    n, lda, ldb = 600, 613, 640
    rng = np.random.default_rng(8)
    A = rng.uniform(0.0, 1.0, (n, n))
    A[n - 64:, :] += 10.0 * np.arange(1, 65)[:, None]
    b = rng.standard_normal((n, 2))
    Ap = np.zeros((lda, n), order="F")
    Ap[:n] = A
    bp = np.full((ldb, 2), 7.0, order="F")
    bp[:n] = b
    dA, dB = ctx.alloc(8 * lda * n), ctx.alloc(8 * ldb * 2)
    try:
        ctx.h2d(dA, Ap)
        ctx.h2d(dB, bp)
        ctx.check(ctx.lib.psx_lu_solve_d(ctx.handle, n, dA, lda, dB, ldb, 2), "lu_solve_d")
        got = np.empty((ldb, 2), order="F")
        ctx.d2h(got, dB)
    finally:
        ctx.free(dA)
        ctx.free(dB)
    x = got[:n]
    eta = np.linalg.norm(A @ x - b) / (np.linalg.norm(A, 2) * np.linalg.norm(x)
                                        + np.linalg.norm(b))
    assert eta < 1e-15
    R = sl.lu_solve(sl.lu_factor(A), b)
    assert np.linalg.norm(x - R) / np.linalg.norm(R) < np.linalg.cond(A) * 1e-15
    assert (got[n:] == 7.0).all()


def _ill_conditioned(n, cond, seed):
    """Q1 diag(logspace(0, -log10 cond)) Q2 + a small identity part: cond(A) ~ cond."""
    rng = np.random.default_rng(seed)
    q1, _ = np.linalg.qr(rng.standard_normal((n, n)))
    q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return (q1 * np.logspace(0.0, -np.log10(cond), n)) @ q2


@pytest.mark.parametrize("n,nrhs", [(300, 37), (700, 700)])
def test_lu_solve_many_rhs_backward_error(ctx, n, nrhs):
    """More than 4 right-hand sides run the blocked substitution (diagonal blocks
    solved by substitution, rocBLAS GEMM updates).  On cond ~ 1e10 the normwise
    backward error stays within 10x of numpy's LAPACK solve (~1e-16)."""
    A = _ill_conditioned(n, 1e10, n)
    B = np.random.default_rng(1).standard_normal((n, nrhs))
    X = ctx.lu_solve(A, B)
    R = np.linalg.solve(A, B)
    eta = lambda Z: np.linalg.norm(A @ Z - B) / (np.linalg.norm(A, 2) * np.linalg.norm(Z)
                                                 + np.linalg.norm(B))
    assert eta(X) < 10 * eta(R) + 1e-16, (eta(X), eta(R))
    assert np.linalg.norm(X - R) / np.linalg.norm(R) < 1e10 * 1e-14


def test_decouple_ill_conditioned(ctx):
    """decouple_covmat(Y, B, B) with cond(B) ~ 1e8: B D B^T reproduces Y within
    10x of numpy's two LAPACK solves.  (On the TT mode-coupling matrix of the
    lmax-2508 covariance workload, cond 2.4e8, rocSOLVER's getrs, whose trsm
    inverts diagonal blocks, reproduced Y to 1.4e-7 against numpy's 2.6e-9;
    the substitution gives 2-4e-9.)"""
    n = 600
    Bm = _ill_conditioned(n, 1e8, 3)
    X = np.random.default_rng(2).standard_normal((n, n))
    Y = X @ X.T / n
    D = ctx.decouple_covmat(Y, Bm, Bm)
    R = np.linalg.solve(Bm, np.linalg.solve(Bm, Y).T).T
    rec = lambda Z: np.linalg.norm(Bm @ Z @ Bm.T - Y) / np.linalg.norm(Y)
    assert rec(D) < 10 * rec(R) + 1e-15, (rec(D), rec(R))
    assert np.linalg.norm(D - R) / np.linalg.norm(R) < 1e8 * 1e-14


def test_decouple_shared_matrix_factored_once(ctx):
    """decouple_covmat(Y, M, M) with one array on both sides factors it once;
    the result equals the call with an equal copy as B2 bit for bit."""
    lmax = 400
    M = ctx.mcm(0, 0, lmax, mask_like_W(lmax + 1, seed=12))
    X = np.random.default_rng(6).standard_normal(M.shape)
    Y = np.asfortranarray(X @ X.T)
    same = ctx.decouple_covmat(Y, M, M)
    apart = ctx.decouple_covmat(Y, M, np.array(M, order="F", copy=True))
    assert np.array_equal(same, apart)
    assert np.linalg.norm(M @ same @ M.T - Y) / np.linalg.norm(Y) < 1e-11


@pytest.mark.parametrize("n,nrhs", [(1, 1), (300, 1), (300, 3), (1700, 2)])
def test_lu_solve_pm(ctx, pkg, n, nrhs):
    """psx_lu_solve_pm: [M++ s M--; s M-- M++] X = B with the 2n x 2n matrix
    assembled on the device equals psx_lu_solve on the matrix assembled on the
    host, bit for bit for s = +1, and for s = -1, where it solves the s = +1
    matrix against D B (D = diag(I, -I)) exactly; both against numpy."""
    rng = np.random.default_rng(n + nrhs)
    Mpp = np.eye(n) * 2.0 + rng.standard_normal((n, n)) / np.sqrt(n)
    Mmm = 0.3 * rng.standard_normal((n, n)) / np.sqrt(n)
    B = rng.standard_normal((2 * n, nrhs))
    for sign in (1, -1):
        A = np.block([[Mpp, sign * Mmm], [sign * Mmm, Mpp]])
        X = ctx.lu_solve_pm(Mpp, Mmm, B, sign)
        assert np.array_equal(X, ctx.lu_solve(A, B)), sign
        R = np.linalg.solve(A, B)
        assert np.linalg.norm(X - R) / np.linalg.norm(R) < 1e-12, sign
    x = ctx.lu_solve_pm(Mpp, Mmm, B[:, 0], -1)
    assert x.shape == (2 * n,) and np.array_equal(x, ctx.lu_solve_pm(Mpp, Mmm, B[:, :1], -1)[:, 0])


def test_lu_solve_pm_arguments(ctx, pkg):
    """A sign other than +-1, blocks of different sizes, a right-hand side of
    the wrong length, and M++ = M-- (a singular block matrix) are refused."""
    A = np.eye(4)
    with pytest.raises(pkg.ArgumentError):
        ctx.lu_solve_pm(A, A, np.ones(8), 0)
    with pytest.raises(pkg.AxisError):
        ctx.lu_solve_pm(A, np.eye(3), np.ones(8))
    with pytest.raises(pkg.AxisError):
        ctx.lu_solve_pm(A, A, np.ones(7))
    with pytest.raises(pkg.SingularException):
        ctx.lu_solve_pm(A, A, np.ones(8), 1)
