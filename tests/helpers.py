"""Shared helpers for the parity tests."""
import numpy as np

# FP64 parity bound between the HIP path and the CPU oracle (relative).
RTOL_FP64 = 1e-12


def rel_err(A, R, floor=1e-6):
    """Max elementwise |A-R|/|R| over entries with |R| >= floor * max|R|."""
    A = np.asarray(A)
    R = np.asarray(R)
    m = np.isfinite(R)
    scale = np.abs(R[m]).max() if m.any() else 0.0
    if scale == 0.0:
        return float(np.abs(A[m]).max()) if m.any() else 0.0
    sel = m & (np.abs(R) >= floor * scale)
    return float((np.abs(A[sel] - R[sel]) / np.abs(R[sel])).max())


def norm_err(A, R):
    m = np.isfinite(R)
    d = np.linalg.norm((np.asarray(A) - np.asarray(R))[m])
    r = np.linalg.norm(np.asarray(R)[m])
    return float(d / r) if r > 0 else float(d)


def mask_like_W(LW, seed=0, cross=False, scale=1.0):
    """Synthetic mask cross-spectrum with the shape of an apodised cut-sky
    mask: W_0 ~ 4 pi f_sky^2, a steep fall-off of many decades and ripples.
    With cross=True some entries are negative (cross-masks)."""
    rng = np.random.default_rng(seed)
    ell = np.arange(LW, dtype=np.float64)
    W = 4 * np.pi * 0.4 * np.exp(-ell / (0.02 * max(LW, 50))) * (1.0 + 0.3 * np.cos(ell / 3.0))
    W += 1e-3 * np.exp(-ell / (0.3 * max(LW, 50))) * rng.uniform(0.5, 1.5, LW)
    W[0] = 4 * np.pi * 0.6
    if cross:
        W *= np.where(rng.uniform(size=LW) < 0.2, -0.5, 1.0)
    return W * scale


def spectra_set(lmax, seed=0):
    """Positive signal-like spectra and noise ratios of length lmax+1."""
    rng = np.random.default_rng(seed)
    ell = np.arange(lmax + 1, dtype=np.float64)
    out = []
    for _ in range(4):
        A = rng.uniform(0.5, 2.0)
        out.append(A * (ell + 10.0) ** -2.5 * np.exp(-(ell / 2000.0) ** 2) * 1e4)
    ratios = [np.sqrt(1.0 + rng.uniform(0.0, 2.0) * (ell / (lmax + 1.0)) ** 2) for _ in range(4)]
    return out, ratios
