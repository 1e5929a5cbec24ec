"""Oracle for the small-shift precision variant (FP64).

Test infrastructure only (see ``oracle/__init__.py``).

The v606 score at an integer lag (ly, lx) is a fixed linear functional of the
packed product Q = T * P (v606:1418-1449):

    ccg(ly, lx) = 1/n^2 * sum_{r,c} Re( Q[r][c] * Phi[(r,c), (ly,lx)] )

(see ``fourier.correlation_functional_unshifted``).  The precision variant
keeps only the coefficients with |k| <= R (a resolution limit) and the lags
|ly|, |lx| <= L, and evaluates the sum exactly here in float64.
"""
from __future__ import annotations

import numpy as np


def disk_indices(n: int, radius: float) -> np.ndarray:
    """Flat packed indices r*(n/2+1) + c of the coefficients with |k| <= radius.

    Frequencies: fy = r (r < n/2) else r - n; fx = c (c < n/2) else -n/2 (the
    packed Nyquist column).  Ascending order."""
    m = n // 2 + 1
    r = np.arange(n)
    fy = np.where(r < n // 2, r, r - n)
    c = np.arange(m)
    fx = np.where(c < n // 2, c, -(n // 2))
    ok = fy[:, None] ** 2 + fx[None, :] ** 2 <= radius * radius
    return np.nonzero(ok.ravel())[0].astype(np.int64)


def lag_grid(half: int):
    """Lags (ly, lx) in row-major order over [-half, half]^2."""
    l = np.arange(-half, half + 1)
    ly, lx = np.meshgrid(l, l, indexing="ij")
    return ly.ravel(), lx.ravel()


def phi_table(n: int, kidx: np.ndarray, half: int) -> np.ndarray:
    """Phi[k, s] (complex128) for the UNshifted packed layout, without the 1/n^2."""
    m = n // 2 + 1
    r, c = kidx // m, kidx % m
    pos = (c + m // 2) % m  # column position after the pre-shift roll
    w = np.where((pos == 0) | (pos == n // 2), 1.0, 2.0)
    ly, lx = lag_grid(half)
    sign = np.where(ly % 2 == 0, 1.0, -1.0)
    ph = 2j * np.pi * (np.outer(r, ly) + np.outer(pos, lx)) / n
    return (w[:, None] * sign[None, :]) * np.exp(ph)


def scores_f64(T: np.ndarray, P: np.ndarray, phi: np.ndarray) -> np.ndarray:
    """scores[a, b, s] = 1024 * sum_k Re(T[a,k] P[b,k] Phi[k,s]) in float64.

    T [A, K], P [B, K] complex (unit-norm over the FULL packed plane, UNshifted)."""
    t = np.asarray(T, np.complex128)
    p = np.asarray(P, np.complex128)
    q = (t[:, None, :] * p[None, :, :]).reshape(-1, t.shape[1])
    return 1024.0 * (q @ np.asarray(phi, np.complex128)).real.reshape(t.shape[0], p.shape[0], -1)


def best_over_lags(scores: np.ndarray):
    """First maximum over the lag axis: (score, index)."""
    k = np.argmax(scores, axis=-1)
    return np.take_along_axis(scores, k[..., None], -1)[..., 0], k


def search_particle(T: np.ndarray, P: np.ndarray, phi: np.ndarray):
    """One particle's small-shift search in FP64.

    T [M, K] (its local directions), P [Psi, K].  Per (direction, psi) pair
    the first maximum over lags; over the pairs (direction-major, as the
    reference flattens scores, v606:1452-1491) the first maximum, and the
    population mean / std of the pair maxima.
    Returns (best pair, best lag, peak, mean, std, pair maxima [M*Psi], pair lags)."""
    s = scores_f64(T, P, phi).reshape(T.shape[0] * P.shape[0], -1)
    v, lag = best_over_lags(s)
    b = int(np.argmax(v))
    return b, int(lag[b]), float(v[b]), float(v.mean()), float(v.std()), v, lag
