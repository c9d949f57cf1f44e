"""ctypes wrapper of the CPU oracle (oracle/libpsx_oracle.so).

TEST INFRASTRUCTURE.  Only tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg may import this module; the product path never does.
The C sources restate the reference algorithm (see psx_oracle.h for what pins
each function).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libpsx_oracle.so")
_lib = None

MCM_00, MCM_02, MCM_PP, MCM_MM = 0, 1, 2, 3
COV_TTTT, COV_EEEE, COV_TTTE, COV_TETE, COV_TEEE, COV_TEEE_PLANCK, COV_TTEE = range(7)
COV_ARITY = {0: (8, 4, 4), 1: (8, 4, 4), 2: (4, 4, 2), 3: (5, 4, 2),
             4: (4, 4, 2), 5: (4, 4, 2), 6: (2, 4, 0)}

_dp = ctypes.POINTER(ctypes.c_double)


def build() -> str:
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.psxo_wigner_family.argtypes = [ctypes.c_int] * 4 + [_dp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.psxo_wigner_family_ld.argtypes = [ctypes.c_int] * 4 + [_dp, _dp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.psxo_mcm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int64, _dp,
                               ctypes.c_int64, ctypes.c_int, ctypes.c_int]
        L.psxo_mcm_rows.argtypes = [ctypes.c_int] * 5 + [_dp, ctypes.c_int64, _dp, ctypes.c_int64,
                                                        ctypes.c_int, ctypes.c_int]
        L.psxo_cov_rows.argtypes = [ctypes.c_int] * 5 + [ctypes.POINTER(_dp)] * 3 + [
            ctypes.c_int64, _dp, ctypes.c_int64, ctypes.c_int]
        L.psxo_cov.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_dp),
                               ctypes.POINTER(_dp), ctypes.POINTER(_dp), ctypes.c_int64, _dp,
                               ctypes.c_int64, ctypes.c_int]
        L.psxo_decouple_covmat.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
        L.psxo_lu_solve.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int]
        L.psxo_quickpol_xi.argtypes = [ctypes.c_int] * 7 + [_dp, ctypes.c_int64, _dp, ctypes.c_int]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def wigner_family(l1: int, l2: int, m1: int, m2: int, precision: str = "double"):
    """Family f(l3) = (l3 l1 l2; -m1-m2 m1 m2); returns (nmin, values)."""
    cap = l1 + l2 + 1
    nmin = ctypes.c_int(0)
    if precision == "double":
        out = np.zeros(cap)
        n = lib().psxo_wigner_family(l1, l2, m1, m2, _ptr(out), cap, ctypes.byref(nmin))
        return nmin.value, out[:n]
    hi = np.zeros(cap)
    lo = np.zeros(cap)
    n = lib().psxo_wigner_family_ld(l1, l2, m1, m2, _ptr(hi), _ptr(lo), cap, ctypes.byref(nmin))
    return nmin.value, hi[:n], lo[:n]


def mcm(kind: int, lmin: int, lmax: int, W: np.ndarray, nthreads: int = 0,
        precision: int = 0) -> np.ndarray:
    """inner_mcm of the given kind; returns the n x n matrix (numpy, row index = l1)."""
    W = np.ascontiguousarray(W, dtype=np.float64)
    n = lmax - lmin + 1
    M = np.zeros((n, n), order="F")
    st = lib().psxo_mcm(kind, lmin, lmax, _ptr(W), W.size, M.ctypes.data_as(_dp), n, nthreads,
                        precision)
    if st != 0:
        raise RuntimeError(f"psxo_mcm status {st}")
    return M


def mcm_rows(kind: int, lmin: int, lmax: int, l1_lo: int, l1_hi: int, W: np.ndarray,
             nthreads: int = 0, precision: int = 0, out: np.ndarray | None = None) -> np.ndarray:
    """Rows l1_lo..l1_hi of the upper triangle (mirrored); other entries NaN."""
    W = np.ascontiguousarray(W, dtype=np.float64)
    n = lmax - lmin + 1
    M = np.full((n, n), np.nan, order="F") if out is None else out
    st = lib().psxo_mcm_rows(kind, lmin, lmax, l1_lo, l1_hi, _ptr(W), W.size,
                             M.ctypes.data_as(_dp), n, nthreads, precision)
    if st != 0:
        raise RuntimeError(f"psxo_mcm_rows status {st}")
    return M


def cov(kind: int, lmin: int, lmax: int, spectra, ratios, W, nthreads: int = 0,
        rows: tuple | None = None, out: np.ndarray | None = None) -> np.ndarray:
    nW, nS, nR = COV_ARITY[kind]
    spectra = [np.ascontiguousarray(s, dtype=np.float64) for s in spectra]
    ratios = [np.ascontiguousarray(r, dtype=np.float64) for r in ratios]
    W = [np.ascontiguousarray(w, dtype=np.float64) for w in W]
    assert len(spectra) == nS and len(ratios) == nR and len(W) == nW
    LW = W[0].size
    assert all(w.size == LW for w in W)
    sp = (_dp * max(nS, 1))(*[_ptr(s) for s in spectra])
    rp = (_dp * max(nR, 1))(*[_ptr(r) for r in ratios])
    wp = (_dp * nW)(*[_ptr(w) for w in W])
    n = lmax - lmin + 1
    lo, hi = rows if rows is not None else (lmin, lmax)
    C = np.full((n, n), np.nan, order="F") if out is None else out
    st = lib().psxo_cov_rows(kind, lmin, lmax, lo, hi, sp, rp, wp, LW, C.ctypes.data_as(_dp), n,
                             nthreads)
    if st != 0:
        raise RuntimeError(f"psxo_cov status {st}")
    return C


def decouple_covmat(Y, B1, B2) -> np.ndarray:
    Y = np.asfortranarray(Y, dtype=np.float64)
    B1 = np.asfortranarray(B1, dtype=np.float64)
    B2 = np.asfortranarray(B2, dtype=np.float64)
    n = Y.shape[0]
    out = np.zeros((n, n), order="F")
    st = lib().psxo_decouple_covmat(n, Y.ctypes.data_as(_dp), B1.ctypes.data_as(_dp),
                                    B2.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    if st != 0:
        raise RuntimeError(f"psxo_decouple_covmat status {st}")
    return out


def lu_solve(A, B) -> np.ndarray:
    A = np.asfortranarray(A, dtype=np.float64)
    B = np.array(B, dtype=np.float64, order="F", copy=True)
    if B.ndim == 1:
        B = B.reshape(-1, 1, order="F")
    n = A.shape[0]
    st = lib().psxo_lu_solve(n, A.ctypes.data_as(_dp), B.ctypes.data_as(_dp), B.shape[1])
    if st != 0:
        raise RuntimeError(f"psxo_lu_solve status {st}")
    return B


def quickpol_xi(nu1: int, nu2: int, s1: int, s2: int, lmax: int, bl: int, bu: int, W,
                band=None, nthreads: int = 0) -> np.ndarray:
    """QuickPol Ξ (src/beam.jl:72-101) in band storage, shape (bl + bu + 1, lmax + 1),
    Fortran order; `band` (in/out) defaults to zeros."""
    W = np.ascontiguousarray(W, dtype=np.float64)
    if band is None:
        band = np.zeros((bl + bu + 1, lmax + 1), order="F")
    if band.shape != (bl + bu + 1, lmax + 1) or band.dtype != np.float64 or \
            not band.flags.f_contiguous:
        raise ValueError("band must be a Fortran-ordered float64 (bl+bu+1, lmax+1) array")
    st = lib().psxo_quickpol_xi(nu1, nu2, s1, s2, lmax, bl, bu, _ptr(W), W.size, _ptr(band),
                                nthreads)
    if st != 0:
        raise RuntimeError(f"psxo_quickpol_xi: status {st}")
    return band
