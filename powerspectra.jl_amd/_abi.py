"""ctypes binding of the C ABI in include/powerspectra_hip.h (libpsx_hip.so).

This is the only way the Python mirror reaches the device: every array goes
through the HIP library.  If the shared library is missing the import fails
loudly -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpsx_hip.so")

# status codes (enum psx_status)
PSX_OK = 0
PSX_ERR_ARGUMENT = 1
PSX_ERR_AXES = 2
PSX_ERR_SINGULAR = 3
PSX_ERR_DOMAIN = 4
PSX_ERR_BUDGET = 5
PSX_ERR_DEVICE = 6
PSX_ERR_NOT_READY = 7

# enum psx_mcm_kind
MCM_00, MCM_02, MCM_PP, MCM_MM = 0, 1, 2, 3
# enum psx_cov_kind
COV_TTTT, COV_EEEE, COV_TTTE, COV_TETE, COV_TEEE, COV_TEEE_PLANCK, COV_TTEE = range(7)
# enum psx_precision
FP64, MIXED, FP64_MFMA = 0, 1, 2
# enum psx_request_type
REQ_MCM, REQ_MCM_PM, REQ_MCM_SET, REQ_COV = 0, 1, 2, 3
# enum psx_sched
SCHED_FIFO, SCHED_DEADLINE = 0, 1


class PsxRequest(ctypes.Structure):
    """struct psx_request of include/powerspectra_hip.h."""
    _fields_ = [("type", ctypes.c_int), ("kind", ctypes.c_int), ("lmin", ctypes.c_int),
                ("lmax", ctypes.c_int), ("LW", ctypes.c_int64),
                ("W", ctypes.c_uint64 * 8), ("spectra", ctypes.c_uint64 * 4),
                ("ratios", ctypes.c_uint64 * 4), ("out", ctypes.c_void_p * 5),
                ("ld", ctypes.c_int64)]


class PowerSpectraError(RuntimeError):
    """Base class of errors reported by the HIP library."""


class ArgumentError(PowerSpectraError, ValueError):
    """Julia ArgumentError: bad spectrum name, kind or size."""


class AxisError(PowerSpectraError, AssertionError):
    """Julia AssertionError on mismatched SpectralArray axes."""


class SingularException(PowerSpectraError, ArithmeticError):
    """LinearAlgebra.SingularException: exact zero pivot in LU."""


class DomainError(PowerSpectraError, ArithmeticError):
    """Julia DomainError: sqrt of a negative spectrum product."""


class BudgetError(PowerSpectraError, MemoryError):
    """Request exceeds the context's HBM budget."""


class DeviceError(PowerSpectraError):
    """HIP, rocBLAS or rocSOLVER failure."""


_ERRORS = {
    PSX_ERR_ARGUMENT: ArgumentError,
    PSX_ERR_AXES: AxisError,
    PSX_ERR_SINGULAR: SingularException,
    PSX_ERR_DOMAIN: DomainError,
    PSX_ERR_BUDGET: BudgetError,
    PSX_ERR_DEVICE: DeviceError,
}

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_int = ctypes.c_int

# name -> (restype, argtypes); mirrors include/powerspectra_hip.h
SIGNATURES = {
    "psx_map2alm": (_int, [_vp, _int, _int, _int, _int, _dp, _dp]),
    "psx_alm2map": (_int, [_vp, _int, _int, _int, _dp, _dp]),
    "psx_lu_factor": (_int, [_vp, _int, _dp, _i64, _dp, _i64, ctypes.POINTER(ctypes.c_int)]),
    "psx_lu_solve_d": (_int, [_vp, _int, _vp, _i64, _vp, _i64, _int]),
    "psx_quickpol_xi": (_int, [_vp] + [_int] * 7 + [_dp, _i64, _dp]),
    "psx_quickpol_xi_d": (_int, [_vp] + [_int] * 7 + [_vp, _i64, _vp]),
    "psx_mask_spectra": (_int, [_vp, _int, _int, _int, _int, _dp, _dp]),
    "psx_alm2cl": (_int, [_vp, _int, _int, ctypes.POINTER(_vp), _dp]),
    "psx_lu_solve_pm": (_int, [_vp, _int, _int, _dp, _dp, _i64, _dp, _i64, _int]),
    "psx_master_decouple": (_int, [_vp, _int, _int, _i64] + [_dp] * 4 + [_dp, _dp]),
    "psx_mcm_binned": (_int, [_vp, _int, _int, _int, _dp, _i64, _dp, _int, _dp, _int, _dp,
                              _i64]),
    "psx_master_spectra": (_int, [_vp, _int, _int, _int, ctypes.POINTER(_vp),
                                  ctypes.POINTER(_vp), _int, _dp, _dp]),
    "psx_map2alm_pol": (_int, [_vp, _int, _int, _int, _int, _dp, _dp]),
    "psx_alm2map_pol": (_int, [_vp, _int, _int, _int, _dp, _dp]),
    "psx_server_create": (_int, [_vp, _int, ctypes.c_size_t, _int, ctypes.POINTER(_vp)]),
    "psx_server_destroy": (_int, [_vp]),
    "psx_server_set_policy": (_int, [_vp, _int, ctypes.c_double]),
    "psx_resident_put": (_int, [_vp, ctypes.c_uint64, _dp, _i64]),
    "psx_resident_evict": (_int, [_vp, ctypes.c_uint64]),
    "psx_submit": (_int, [_vp, ctypes.POINTER(PsxRequest), ctypes.POINTER(ctypes.c_uint64)]),
    "psx_poll": (_int, [_vp, ctypes.c_uint64, ctypes.POINTER(_int)]),
    "psx_wait": (_int, [_vp, ctypes.c_uint64, ctypes.POINTER(_int), _dp]),
    "psx_fetch": (_int, [_vp, ctypes.c_uint64, ctypes.POINTER(_vp), _i64]),
    "psx_release": (_int, [_vp, ctypes.c_uint64]),
    "psx_server_last_error": (ctypes.c_char_p, [_vp]),
    "psx_cov_arity": (_int, [_int, ctypes.POINTER(_int), ctypes.POINTER(_int), ctypes.POINTER(_int)]),
    "psx_ctx_create": (_int, [_int, ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "psx_ctx_destroy": (_int, [_vp]),
    "psx_ctx_synchronize": (_int, [_vp]),
    "psx_ctx_last_error": (ctypes.c_char_p, [_vp]),
    "psx_ctx_bytes_in_use": (ctypes.c_size_t, [_vp]),
    "psx_ctx_set_timing": (_int, [_vp, _int]),
    "psx_ctx_kernel_time": (_int, [_vp, ctypes.c_char_p, _int, _dp, ctypes.POINTER(_i64)]),
    "psx_dev_alloc": (_int, [_vp, ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "psx_dev_free": (_int, [_vp, _vp]),
    "psx_memcpy_h2d": (_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    "psx_memcpy_d2h": (_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    "psx_mcm": (_int, [_vp, _int, _int, _int, _dp, _i64, _dp, _i64]),
    "psx_mcm_d": (_int, [_vp, _int, _int, _int, _vp, _i64, _vp, _i64]),
    "psx_mcm_d_noalloc": (_int, [_vp, _int, _int, _int, _vp, _i64, _vp, _i64]),
    "psx_mcm_pm": (_int, [_vp, _int, _int, _dp, _i64, _dp, _dp, _i64]),
    "psx_mcm_pm_d": (_int, [_vp, _int, _int, _vp, _i64, _vp, _vp, _i64]),
    "psx_mcm_set": (_int, [_vp, _int, _int, _i64] + [_dp] * 4 + [_dp] * 5 + [_i64]),
    "psx_mcm_set_d": (_int, [_vp, _int, _int, _i64] + [_vp] * 4 + [_vp] * 5 + [_i64]),
    "psx_cov": (_int, [_vp, _int, _int, _int, ctypes.POINTER(_dp), ctypes.POINTER(_dp),
                       ctypes.POINTER(_dp), _i64, _int, _dp, _i64]),
    "psx_cov_d": (_int, [_vp, _int, _int, _int, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                         ctypes.POINTER(_vp), _i64, _int, _vp, _i64]),
    "psx_decouple_covmat": (_int, [_vp, _int, _dp, _dp, _dp, _dp]),
    "psx_mcm_batch": (_int, [_vp, _int, _int, _int, _int, ctypes.POINTER(_dp), _i64, _int,
                             ctypes.POINTER(_dp), _i64]),
    "psx_mcm_batch_d": (_int, [_vp, _int, _int, _int, _int, ctypes.POINTER(_vp), _i64, _int,
                               ctypes.POINTER(_vp), _i64]),
    "psx_cov_batch": (_int, [_vp, _int, _int, _int, _int, ctypes.POINTER(_dp),
                             ctypes.POINTER(_dp), ctypes.POINTER(_dp), _i64, _int,
                             ctypes.POINTER(_dp), _i64]),
    "psx_cov_batch_d": (_int, [_vp, _int, _int, _int, _int, ctypes.POINTER(_vp),
                               ctypes.POINTER(_vp), ctypes.POINTER(_vp), _i64, _int,
                               ctypes.POINTER(_vp), _i64]),
    "psx_lu_solve": (_int, [_vp, _int, _dp, _i64, _dp, _i64, _int]),
}

_lib = None
_lib_lock = threading.Lock()


def lib():
    """Load libpsx_hip.so (raises if it has not been built)."""
    global _lib
    with _lib_lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(
                    f"{LIB_PATH} is missing: build it with powerspectra.jl_amd/build.py "
                    "(the HIP library is the only compute path)")
            L = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _block_size(lmin: int, lmax: int) -> int:
    if lmin < 0 or lmax < lmin:
        raise ArgumentError(f"need 0 <= lmin <= lmax (got lmin={lmin}, lmax={lmax})")
    return lmax - lmin + 1


class Context:
    """One HIP device context (psx_ctx): a stream, an HBM budget, caches."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        L = lib()
        h = _vp()
        st = L.psx_ctx_create(device, hbm_budget, ctypes.byref(h))
        if st != PSX_OK:
            raise _ERRORS.get(st, PowerSpectraError)(f"psx_ctx_create failed with status {st}")
        self._h = h
        self.device = device
        self.lib = L

    @property
    def handle(self):
        return self._h

    def check(self, st: int, what: str = ""):
        if st != PSX_OK:
            msg = self.lib.psx_ctx_last_error(self._h)
            msg = msg.decode() if msg else ""
            raise _ERRORS.get(st, PowerSpectraError)(f"{what}: {msg} (status {st})")

    def close(self):
        """Destroy the context.  Refused (ArgumentError, the context stays
        usable) while a Server created on it is still open."""
        if getattr(self, "_h", None):
            self.check(self.lib.psx_ctx_destroy(self._h), "psx_ctx_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- utilities -------------------------------------------------------
    def synchronize(self):
        self.check(self.lib.psx_ctx_synchronize(self._h), "synchronize")

    def bytes_in_use(self) -> int:
        return int(self.lib.psx_ctx_bytes_in_use(self._h))

    def set_timing(self, on: bool):
        self.check(self.lib.psx_ctx_set_timing(self._h, int(bool(on))), "set_timing")

    def kernel_time(self, name: str, reset: bool = False):
        ms = ctypes.c_double(0)
        n = _i64(0)
        self.check(self.lib.psx_ctx_kernel_time(self._h, name.encode(), int(reset),
                                                ctypes.byref(ms), ctypes.byref(n)), "kernel_time")
        return ms.value, n.value

    def alloc(self, nbytes: int) -> int:
        p = _vp()
        self.check(self.lib.psx_dev_alloc(self._h, nbytes, ctypes.byref(p)), "dev_alloc")
        return p.value

    def free(self, p: int):
        self.check(self.lib.psx_dev_free(self._h, _vp(p)), "dev_free")

    def h2d(self, dst: int, src: np.ndarray):
        """Copy the bytes of src as laid out (C- or Fortran-contiguous arrays are
        sent as they are; anything else is made C-contiguous first)."""
        src = np.asarray(src)
        if not (src.flags.c_contiguous or src.flags.f_contiguous):
            src = np.ascontiguousarray(src)
        self.check(self.lib.psx_memcpy_h2d(self._h, _vp(dst), src.ctypes.data_as(_vp), src.nbytes),
                   "memcpy_h2d")

    def d2h(self, dst: np.ndarray, src: int):
        assert dst.flags.c_contiguous or dst.flags.f_contiguous
        self.check(self.lib.psx_memcpy_d2h(self._h, dst.ctypes.data_as(_vp), _vp(src), dst.nbytes),
                   "memcpy_d2h")

    # --- compute ---------------------------------------------------------
    def mcm(self, kind: int, lmin: int, lmax: int, W: np.ndarray, out: np.ndarray | None = None):
        W = np.ascontiguousarray(W, dtype=np.float64)
        n = _block_size(lmin, lmax)
        if out is None:
            out = np.empty((n, n), order="F")
        self.check(self.lib.psx_mcm(self._h, kind, lmin, lmax, dptr(W), W.size, dptr(out), n),
                   "psx_mcm")
        return out

    def mcm_binned(self, kind: int, lmin: int, lmax: int, W: np.ndarray, P: np.ndarray,
                   Q: np.ndarray) -> np.ndarray:
        """P M Q for one mode-coupling block M over l = lmin..lmax: P is
        (nbl, n), Q is (n, nbr); M stays on the device."""
        W = np.ascontiguousarray(W, dtype=np.float64)
        n = _block_size(lmin, lmax)
        P = np.asfortranarray(P, dtype=np.float64)
        Q = np.asfortranarray(Q, dtype=np.float64)
        if P.ndim != 2 or Q.ndim != 2 or P.shape[1] != n or Q.shape[0] != n:
            raise AxisError(f"binning matrices must be (nbl, {n}) and ({n}, nbr)")
        out = np.empty((P.shape[0], Q.shape[1]), order="F")
        self.check(self.lib.psx_mcm_binned(self._h, kind, lmin, lmax, dptr(W), W.size, dptr(P),
                                           P.shape[0], dptr(Q), Q.shape[1], dptr(out),
                                           P.shape[0]), "psx_mcm_binned")
        return out

    def mcm_pm(self, lmin: int, lmax: int, W: np.ndarray):
        W = np.ascontiguousarray(W, dtype=np.float64)
        n = _block_size(lmin, lmax)
        P = np.empty((n, n), order="F")
        M = np.empty((n, n), order="F")
        self.check(self.lib.psx_mcm_pm(self._h, lmin, lmax, dptr(W), W.size, dptr(P), dptr(M), n),
                   "psx_mcm_pm")
        return P, M

    def mcm_set(self, lmin: int, lmax: int, W_TT, W_TE, W_ET, W_PP, out=None):
        """The five blocks (TT, TE, ET, ++, --) in one pass; `out` may hold five
        preallocated n x n Fortran-ordered float64 arrays (None entries skip)."""
        Ws = [None if w is None else np.ascontiguousarray(w, dtype=np.float64)
              for w in (W_TT, W_TE, W_ET, W_PP)]
        LW = next(w.size for w in Ws if w is not None)
        if any(w is not None and w.size != LW for w in Ws):
            raise ArgumentError("all W spectra must have the same length")
        n = _block_size(lmin, lmax)
        want = [Ws[0] is not None, Ws[1] is not None, Ws[2] is not None,
                Ws[3] is not None, Ws[3] is not None]
        if out is None:
            outs = [np.empty((n, n), order="F") if w else None for w in want]
        else:
            outs = list(out)
            for o, w in zip(outs, want):
                if w and (o is None or o.shape != (n, n) or o.dtype != np.float64
                          or not o.flags.f_contiguous):
                    raise ArgumentError("out: five n x n Fortran-ordered float64 arrays expected")
            outs = [o if w else None for o, w in zip(outs, want)]
        nullp = ctypes.cast(None, _dp)
        args = [nullp if w is None else dptr(w) for w in Ws]
        oargs = [nullp if o is None else dptr(o) for o in outs]
        self.check(self.lib.psx_mcm_set(self._h, lmin, lmax, LW, *args, *oargs, n), "psx_mcm_set")
        return outs

    def cov(self, kind: int, lmin: int, lmax: int, spectra, ratios, W, precision: int = FP64):
        nW, nS, nR = cov_arity(kind)
        _block_size(lmin, lmax)
        if len(W) != nW or len(spectra) != nS or len(ratios) != nR:
            raise ArgumentError(f"covariance kind {kind} takes {nW} W, {nS} spectra, {nR} ratios")
        spectra = [np.ascontiguousarray(s, dtype=np.float64) for s in spectra]
        ratios = [np.ascontiguousarray(r, dtype=np.float64) for r in ratios]
        W = [np.ascontiguousarray(w, dtype=np.float64) for w in W]
        for s in spectra + ratios:
            if s.size < lmax + 1:
                raise ArgumentError("spectra and noise ratios need at least lmax+1 entries")
        LW = W[0].size
        if any(w.size != LW for w in W):
            raise ArgumentError("all W spectra must have the same length")
        n = lmax - lmin + 1
        out = np.empty((n, n), order="F")
        sp = (_dp * 4)(*[dptr(s) for s in spectra])
        rp = (_dp * 4)(*([dptr(r) for r in ratios] + [ctypes.cast(None, _dp)] * (4 - nR)))
        wp = (_dp * 8)(*([dptr(w) for w in W] + [ctypes.cast(None, _dp)] * (8 - nW)))
        self.check(self.lib.psx_cov(self._h, kind, lmin, lmax, sp, rp, wp, LW, precision,
                                    dptr(out), n), "psx_cov")
        return out

    def mcm_batch(self, kind: int, lmin: int, lmax: int, Ws, precision: int = FP64):
        """K mode-coupling blocks of one kind from K W spectra in one pass
        (psx_mcm_batch); returns K n x n Fortran-ordered arrays."""
        Ws = [np.ascontiguousarray(w, dtype=np.float64) for w in Ws]
        if not Ws:
            raise ArgumentError("at least one W spectrum")
        LW = Ws[0].size
        if any(w.size != LW for w in Ws):
            raise ArgumentError("all W spectra must have the same length")
        n = _block_size(lmin, lmax)
        outs = [np.empty((n, n), order="F") for _ in Ws]
        wp = (_dp * len(Ws))(*[dptr(w) for w in Ws])
        op = (_dp * len(Ws))(*[dptr(o) for o in outs])
        self.check(self.lib.psx_mcm_batch(self._h, kind, lmin, lmax, len(Ws), wp, LW, precision,
                                          op, n), "psx_mcm_batch")
        return outs

    def cov_batch(self, kind: int, lmin: int, lmax: int, spectra, ratios, W,
                  precision: int = FP64):
        """Covariance blocks of one kind (psx_cov_batch): spectra[b], ratios[b]
        and W[b] are block b's arguments in the reference loop's order."""
        nW, nS, nR = cov_arity(kind)
        nb = len(W)
        _block_size(lmin, lmax)
        if len(spectra) != nb or len(ratios) != nb:
            raise ArgumentError("spectra, ratios and W must list the same blocks")
        sp, rt, ws = [], [], []
        for b in range(nb):
            if len(W[b]) != nW or len(spectra[b]) != nS or len(ratios[b]) != nR:
                raise ArgumentError(f"covariance kind {kind} takes {nW} W, {nS} spectra, "
                                    f"{nR} ratios per block")
            sp += [np.ascontiguousarray(x, dtype=np.float64) for x in spectra[b]]
            rt += [np.ascontiguousarray(x, dtype=np.float64) for x in ratios[b]]
            ws += [np.ascontiguousarray(x, dtype=np.float64) for x in W[b]]
        for x in sp + rt:
            if x.size < lmax + 1:
                raise ArgumentError("spectra and noise ratios need at least lmax+1 entries")
        LW = ws[0].size
        if any(w.size != LW for w in ws):
            raise ArgumentError("all W spectra must have the same length")
        n = lmax - lmin + 1
        outs = [np.empty((n, n), order="F") for _ in range(nb)]
        spp = (_dp * max(len(sp), 1))(*[dptr(x) for x in sp])
        rtp = (_dp * max(len(rt), 1))(*[dptr(x) for x in rt])
        wsp = (_dp * len(ws))(*[dptr(x) for x in ws])
        op = (_dp * nb)(*[dptr(o) for o in outs])
        self.check(self.lib.psx_cov_batch(self._h, kind, lmin, lmax, nb, spp, rtp, wsp, LW,
                                          precision, op, n), "psx_cov_batch")
        return outs

    def map2alm(self, maps, nside: int, lmax: int, niter: int = 3) -> np.ndarray:
        """a_lm (complex, m-major, shape (nmaps, nalm)) of ring-ordered maps."""
        M = np.ascontiguousarray(np.atleast_2d(maps), dtype=np.float64)
        if M.shape[1] != 12 * nside * nside:
            raise ArgumentError("maps must have 12 nside^2 pixels")
        nalm = (lmax + 1) * (lmax + 2) // 2
        out = np.empty((M.shape[0], nalm), dtype=np.complex128)
        self.check(self.lib.psx_map2alm(self._h, nside, lmax, niter, M.shape[0], dptr(M),
                                        out.ctypes.data_as(_dp)), "psx_map2alm")
        return out

    def alm2map(self, alms, nside: int, lmax: int) -> np.ndarray:
        A = np.ascontiguousarray(np.atleast_2d(alms), dtype=np.complex128)
        if A.shape[1] != (lmax + 1) * (lmax + 2) // 2:
            raise ArgumentError("alm length does not match lmax")
        out = np.empty((A.shape[0], 12 * nside * nside))
        self.check(self.lib.psx_alm2map(self._h, nside, lmax, A.shape[0],
                                        A.ctypes.data_as(_dp), dptr(out)), "psx_alm2map")
        return out

    def mask_spectra(self, maps, nside: int, lmax: int, niter: int = 3) -> np.ndarray:
        """cl[i, j, l] = alm2cl(map2alm(m_i), map2alm(m_j))[l] of 1..8 ring-ordered
        maps, alms kept on the device."""
        M = np.ascontiguousarray(np.atleast_2d(maps), dtype=np.float64)
        if M.shape[1] != 12 * nside * nside:
            raise ArgumentError("maps must have 12 nside^2 pixels")
        n = M.shape[0]
        out = np.empty((n, n, lmax + 1))
        self.check(self.lib.psx_mask_spectra(self._h, nside, lmax, niter, n, dptr(M),
                                             dptr(out)), "psx_mask_spectra")
        return out

    def alm2cl(self, alms, lmax: int) -> np.ndarray:
        """cl[i, j, l] = alm2cl(a_i, a_j)[l] of 1..8 coefficient sets (complex,
        m-major, nalm each), summed on the device; the sets are not copied on
        the host when they are already contiguous complex128."""
        nalm = (lmax + 1) * (lmax + 2) // 2
        sets = [np.ascontiguousarray(a, dtype=np.complex128) for a in alms]
        for a in sets:
            if a.ndim != 1 or a.shape[0] != nalm:
                raise ArgumentError("alm length does not match lmax")
        n = len(sets)
        ptrs = (_vp * max(n, 1))(*[a.ctypes.data for a in sets])
        out = np.empty((n, n, lmax + 1))
        self.check(self.lib.psx_alm2cl(self._h, lmax, n, ptrs, dptr(out)), "psx_alm2cl")
        return out

    def master_spectra(self, maps, masks, nside: int, lmax: int, niter: int = 3,
                       already_masked: bool = False):
        """(pcl[x, y, l], W[a, b, l]): the pseudo-spectra of map 1's x in (T, E, B)
        with map 2's y, and the spectra of mask a in (T, P) of map 1 with mask b
        of map 2, from maps = (I1, Q1, U1, I2, Q2, U2) and masks = (T1, P1, T2,
        P2), all on the device (psx_master_spectra)."""
        npix = 12 * nside * nside
        M = [np.ascontiguousarray(m, dtype=np.float64) for m in maps]
        K = [np.ascontiguousarray(m, dtype=np.float64) for m in masks]
        if len(M) != 6 or len(K) != 4:
            raise ArgumentError("six map fields and four masks are required")
        if any(a.ndim != 1 or a.size != npix for a in M + K):
            raise ArgumentError("maps and masks must have 12 nside^2 pixels")
        mp = (_vp * 6)(*[a.ctypes.data for a in M])
        kp = (_vp * 4)(*[a.ctypes.data for a in K])
        pcl = np.empty((3, 3, lmax + 1))
        W = np.empty((2, 2, lmax + 1))
        self.check(self.lib.psx_master_spectra(self._h, nside, lmax, niter, mp, kp,
                                               int(bool(already_masked)), dptr(pcl), dptr(W)),
                   "psx_master_spectra")
        return pcl, W

    def map2alm_pol(self, iqu, nside: int, lmax: int, niter: int = 3) -> np.ndarray:
        """(T, E, B) a_lm, shape (nmaps, 3, nalm), of (I, Q, U) ring-ordered
        maps given as (3, npix) or (nmaps, 3, npix)."""
        M = np.asarray(iqu, dtype=np.float64)
        if M.ndim == 2:
            M = M[None]
        M = np.ascontiguousarray(M)
        if M.ndim != 3 or M.shape[1] != 3 or M.shape[2] != 12 * nside * nside:
            raise ArgumentError("maps must be (nmaps, 3, 12 nside^2)")
        nalm = (lmax + 1) * (lmax + 2) // 2
        out = np.empty((M.shape[0], 3, nalm), dtype=np.complex128)
        self.check(self.lib.psx_map2alm_pol(self._h, nside, lmax, niter, M.shape[0], dptr(M),
                                            out.ctypes.data_as(_dp)), "psx_map2alm_pol")
        return out

    def alm2map_pol(self, teb, nside: int, lmax: int) -> np.ndarray:
        """(I, Q, U) maps, shape (nmaps, 3, npix), of (T, E, B) a_lm given as
        (3, nalm) or (nmaps, 3, nalm)."""
        A = np.asarray(teb, dtype=np.complex128)
        if A.ndim == 2:
            A = A[None]
        A = np.ascontiguousarray(A)
        if A.ndim != 3 or A.shape[1] != 3 or A.shape[2] != (lmax + 1) * (lmax + 2) // 2:
            raise ArgumentError("alms must be (nmaps, 3, nalm) with nalm matching lmax")
        out = np.empty((A.shape[0], 3, 12 * nside * nside))
        self.check(self.lib.psx_alm2map_pol(self._h, nside, lmax, A.shape[0],
                                            A.ctypes.data_as(_dp), dptr(out)), "psx_alm2map_pol")
        return out

    def quickpol_xi(self, nu1: int, nu2: int, s1: int, s2: int, lmax: int, bl: int, bu: int,
                    W, band=None) -> np.ndarray:
        """QuickPol Ξ in band storage (bl + bu + 1, lmax + 1), Fortran order;
        `band` is read and written (zeros if not given)."""
        W = np.ascontiguousarray(W, dtype=np.float64).ravel()
        ldb = bl + bu + 1
        if band is None:
            band = np.zeros((ldb, lmax + 1), order="F")
        if band.shape != (ldb, lmax + 1) or band.dtype != np.float64 or \
                not band.flags.f_contiguous:
            raise AxisError("band must be a Fortran-ordered float64 (bl+bu+1, lmax+1) array")
        self.check(self.lib.psx_quickpol_xi(self._h, nu1, nu2, s1, s2, lmax, bl, bu, dptr(W),
                                            W.size, dptr(band)), "psx_quickpol_xi")
        return band

    def lu_factor(self, A):
        """(LU, ipiv) of a square matrix: LAPACK getrf layout, ipiv 1-based."""
        A = np.asfortranarray(A, dtype=np.float64)
        n = A.shape[0]
        if A.shape != (n, n):
            raise AxisError("lu_factor needs a square matrix")
        LU = np.empty((n, n), order="F")
        ipiv = np.empty(n, dtype=np.int32)
        self.check(self.lib.psx_lu_factor(self._h, n, dptr(A), n, dptr(LU), n,
                                          ipiv.ctypes.data_as(ctypes.POINTER(ctypes.c_int))),
                   "psx_lu_factor")
        return LU, ipiv

    def decouple_covmat(self, Y, B1, B2):
        """B1^{-1} Y B2^{-T}; one array passed as both B1 and B2 reaches the
        library as one pointer, so it is factored once."""
        Y = np.asfortranarray(Y, dtype=np.float64)
        shared = B2 is B1
        B1 = np.asfortranarray(B1, dtype=np.float64)
        B2 = B1 if shared else np.asfortranarray(B2, dtype=np.float64)
        n = Y.shape[0]
        if Y.shape != (n, n) or B1.shape != (n, n) or B2.shape != (n, n):
            raise AxisError("decouple_covmat needs three n x n matrices")
        out = np.empty((n, n), order="F")
        self.check(self.lib.psx_decouple_covmat(self._h, n, dptr(Y), dptr(B1), dptr(B2),
                                                dptr(out)), "psx_decouple_covmat")
        return out

    def lu_solve(self, A, B):
        A = np.asfortranarray(A, dtype=np.float64)
        n = A.shape[0]
        if A.shape != (n, n):
            raise AxisError("lu_solve needs a square matrix")
        vec = np.ndim(B) == 1
        X = np.array(B, dtype=np.float64, order="F", copy=True)
        if vec:
            X = X.reshape(-1, 1, order="F")
        if X.shape[0] != n:
            raise AxisError("right-hand side length does not match the matrix")
        self.check(self.lib.psx_lu_solve(self._h, n, dptr(A), n, dptr(X), n, X.shape[1]),
                   "psx_lu_solve")
        return X[:, 0] if vec else X

    def master_decouple(self, lmin: int, lmax: int, W_TT, W_TE, W_ET, W_PP, pcl):
        """The nine decoupled spectra (TT, TE, TB, ET, BT, EE, BB, EB, BE) from
        the four mask cross-spectra and the nine pseudo-spectra (rows of pcl,
        l = lmin..lmax): the five blocks and the four solves stay on the device
        (psx_master_decouple).  Returns a 9 x n array."""
        Ws = [np.ascontiguousarray(w, dtype=np.float64) for w in (W_TT, W_TE, W_ET, W_PP)]
        LW = Ws[0].size
        if any(w.size != LW for w in Ws):
            raise ArgumentError("all W spectra must have the same length")
        n = _block_size(lmin, lmax)
        P = np.ascontiguousarray(pcl, dtype=np.float64)
        if P.shape != (9, n):
            raise AxisError("pcl: nine pseudo-spectra of lmax - lmin + 1 entries expected")
        X = np.empty((9, n))
        self.check(self.lib.psx_master_decouple(self._h, lmin, lmax, LW, *[dptr(w) for w in Ws],
                                                dptr(P), dptr(X)), "psx_master_decouple")
        return X

    def lu_solve_pm(self, Mpp, Mmm, B, sign: int = 1):
        """[M++ s M--; s M-- M++] \\ B with s = sign (+1: EE/BB, -1: EB/BE), the
        2n x 2n matrix assembled on the device (psx_lu_solve_pm)."""
        Mpp = np.asfortranarray(Mpp, dtype=np.float64)
        Mmm = np.asfortranarray(Mmm, dtype=np.float64)
        n = Mpp.shape[0]
        if Mpp.shape != (n, n) or Mmm.shape != (n, n):
            raise AxisError("lu_solve_pm needs two square blocks of one size")
        vec = np.ndim(B) == 1
        X = np.array(B, dtype=np.float64, order="F", copy=True)
        if vec:
            X = X.reshape(-1, 1, order="F")
        if X.shape[0] != 2 * n:
            raise AxisError("right-hand side length is not twice the block size")
        self.check(self.lib.psx_lu_solve_pm(self._h, n, int(sign), dptr(Mpp), dptr(Mmm), n,
                                            dptr(X), 2 * n, X.shape[1]), "psx_lu_solve_pm")
        return X[:, 0] if vec else X


class Server:
    """Persistent request loop on one context (psx_server): `slots` requests
    run concurrently on the device, admitted in arrival order as slots and
    arena memory free up; spectra are resident under integer keys."""

    _NOUT = {REQ_MCM: 1, REQ_MCM_PM: 2, REQ_MCM_SET: 5, REQ_COV: 1}

    def __init__(self, ctx: "Context", slots: int = 4, arena_bytes: int = 16 << 30,
                 lmax_max: int = 4095):
        self.ctx = ctx
        self.lib = ctx.lib
        h = _vp()
        st = self.lib.psx_server_create(ctx.handle, slots, arena_bytes, lmax_max, ctypes.byref(h))
        self._check(st, "psx_server_create", None)
        self._h = h
        self._pending = {}   # ticket -> (lmin, lmax, nout, host outputs)

    def _check(self, st, what, h="self"):
        if st != PSX_OK:
            msg = self.lib.psx_server_last_error(self._h if h == "self" else None)
            msg = msg.decode() if msg else ""
            raise _ERRORS.get(st, PowerSpectraError)(f"{what}: {msg} (status {st})")

    @property
    def handle(self):
        return self._h

    def set_policy(self, policy: int, slack: float = 2.0):
        self._check(self.lib.psx_server_set_policy(self._h, policy, slack), "psx_server_set_policy")

    def put(self, key: int, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        self._check(self.lib.psx_resident_put(self._h, key, dptr(v), v.size), "psx_resident_put")

    def evict(self, key: int):
        self._check(self.lib.psx_resident_evict(self._h, key), "psx_resident_evict")

    def submit(self, rtype: int, kind: int = 0, lmin: int = 0, lmax: int = 0, LW: int = 0,
               W=(), spectra=(), ratios=(), out=None) -> int:
        """Enqueue a request; `out` = list of n x n Fortran float64 arrays to
        receive the result, or None to keep it on the device (psx_fetch)."""
        n = _block_size(lmin, lmax)
        r = PsxRequest()
        r.type, r.kind, r.lmin, r.lmax = rtype, kind, lmin, lmax
        r.LW = LW or lmax + 1
        for i, k in enumerate(W):
            r.W[i] = k
        for i, k in enumerate(spectra):
            r.spectra[i] = k
        for i, k in enumerate(ratios):
            r.ratios[i] = k
        nout = self._NOUT.get(rtype, 1)
        if out is not None:
            if len(out) != nout:
                raise ArgumentError(f"request type {rtype} has {nout} outputs")
            for i, o in enumerate(out):
                if o.shape != (n, n) or o.dtype != np.float64 or not o.flags.f_contiguous:
                    raise ArgumentError("outputs must be n x n Fortran-ordered float64")
                r.out[i] = o.ctypes.data
            r.ld = n
        t = ctypes.c_uint64(0)
        self._check(self.lib.psx_submit(self._h, ctypes.byref(r), ctypes.byref(t)), "psx_submit")
        self._pending[t.value] = (lmin, lmax, nout, out)
        return t.value

    def poll(self, ticket: int) -> bool:
        d = _int(0)
        self._check(self.lib.psx_poll(self._h, ticket, ctypes.byref(d)), "psx_poll")
        return bool(d.value)

    def wait(self, ticket: int, raise_on_error: bool = True):
        """Block until done; returns (status, (queued s, running s, total s))."""
        st = _int(0)
        times = (ctypes.c_double * 3)()
        self._check(self.lib.psx_wait(self._h, ticket, ctypes.byref(st),
                                      ctypes.cast(times, _dp)), "psx_wait")
        if st.value != PSX_OK and raise_on_error:
            self._check(st.value, f"request {ticket}")
        return st.value, tuple(times)

    def fetch(self, ticket: int):
        """Copy a device-held result to new host arrays."""
        if ticket not in self._pending:
            raise ArgumentError(f"psx_fetch: unknown ticket {ticket}")
        lmin, lmax, nout, _ = self._pending[ticket]
        n = lmax - lmin + 1
        outs = [np.empty((n, n), order="F") for _ in range(nout)]
        arr = (_vp * nout)(*[o.ctypes.data for o in outs])
        self._check(self.lib.psx_fetch(self._h, ticket, arr, n), "psx_fetch")
        return outs

    def release(self, ticket: int):
        self._check(self.lib.psx_release(self._h, ticket), "psx_release")
        self._pending.pop(ticket, None)

    def close(self):
        if getattr(self, "_h", None):
            self.lib.psx_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cov_arity(kind: int):
    nW, nS, nR = _int(0), _int(0), _int(0)
    st = lib().psx_cov_arity(kind, ctypes.byref(nW), ctypes.byref(nS), ctypes.byref(nR))
    if st != PSX_OK:
        raise ArgumentError(f"unknown covariance kind {kind}")
    return nW.value, nS.value, nR.value


_default_ctx = None
_default_lock = threading.Lock()


def default_context() -> Context:
    """Process-wide context on device 0 (created on first use)."""
    global _default_ctx
    with _default_lock:
        if _default_ctx is None:
            _default_ctx = Context(int(os.environ.get("PSX_DEVICE", "0")))
    return _default_ctx
