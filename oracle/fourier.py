"""Oracle: packed Fourier layout, pre-shift and the v606 correlation operator.

Test infrastructure only (see ``oracle/__init__.py``).
"""
from __future__ import annotations

import numpy as np


def packed_row_index(n: int) -> np.ndarray:
    """v606:505-511 centered_row_to_packed: [n/2..n-1, 0..n/2-1]."""
    h = n // 2
    return np.concatenate((np.arange(h, n), np.arange(0, h)))


def packed_col_index(n: int) -> np.ndarray:
    """v606:512-517 centered_col_to_packed: [n/2..n-1, 0]."""
    h = n // 2
    return np.concatenate((np.arange(h, n), np.zeros(1, dtype=np.int64)))


def centered_full_to_packed(centered: np.ndarray) -> np.ndarray:
    """v606:531-533: gather rows then columns of a centred (fftshifted) plane."""
    n = centered.shape[-1]
    p = np.take(centered, packed_row_index(n), axis=-2)
    return np.take(p, packed_col_index(n), axis=-1)


def pre_shift(packed: np.ndarray) -> np.ndarray:
    """v606:535-537: fftshift over the last two axes (rows by n/2, cols by (n/2+1)//2)."""
    return np.fft.fftshift(packed, axes=(-2, -1))


def crop_indices(n: int, crop: int) -> np.ndarray:
    """v606:518-529 unshifted_ccg_crop_indices: [n-C/2..n-1, 0..C/2-1]."""
    h = crop // 2
    return np.concatenate((np.arange(n - h, n), np.arange(0, h)))


def correlate(templates_pre, particles_pre, crop: int, dtype=np.float32):
    """v606:1418-1449: for every (template a, psi b):

    ccg = irfft2(T_a * P_b, s=(n, n)); crop rows/cols at crop_indices;
    score = max * 1024 * n^2; y = argmax // C; x = argmax % C.
    Returns (scores [A,B], y [A,B], x [A,B]).
    """
    t = np.asarray(templates_pre)
    p = np.asarray(particles_pre)
    n = t.shape[-2]
    cdt = np.complex64 if dtype == np.float32 else np.complex128
    t = t.astype(cdt)
    p = p.astype(cdt)
    idx = crop_indices(n, crop)
    A, B = t.shape[0], p.shape[0]
    scores = np.empty((A, B), dtype)
    ys = np.empty((A, B), np.int64)
    xs = np.empty((A, B), np.int64)
    mult = dtype(1024 * n * n)
    for a in range(A):
        prod = t[a][None] * p
        ccg = np.fft.irfft2(prod, s=(n, n), axes=(-2, -1)).astype(dtype)
        cr = ccg[:, idx][:, :, idx].reshape(B, -1)
        k = np.argmax(cr, axis=1)  # first maximum, as torch.max on CPU
        scores[a] = cr[np.arange(B), k] * mult
        ys[a] = k // crop
        xs[a] = k % crop
    return scores, ys, xs


def correlation_functional(n: int, crop: int) -> np.ndarray:
    """The v606 lag map as an explicit linear functional (independent derivation).

    For pre-shifted operands Q = T~ * P~ (packed rfft layout, rows and columns
    rolled by fftshift), irfft2 followed by the corner crop equals

        ccg[i, j] = (1/n^2) * sum_{r,c} w_c * Re(Q[r, c] * e^{2 pi i (r*ly_i + c*lx_j)/n})

    with ly/lx the crop lags and w_c = 1 for c in {0, n/2}, else 2 (the
    complex-to-real rule: the imaginary parts of columns 0 and n/2 are
    dropped after the row transform).  Returns Phi with shape
    [n, n/2+1, C, C] (complex) so that ccg = Re(einsum('rc,rcij', Q, Phi)).
    """
    m = n // 2 + 1
    lag = crop_indices(n, crop)
    w = np.full(m, 2.0)
    w[0] = 1.0
    w[n // 2] = 1.0
    r = np.arange(n)
    c = np.arange(m)
    er = np.exp(2j * np.pi * np.outer(r, lag) / n)            # [n, C]
    ec = np.exp(2j * np.pi * np.outer(c, lag) / n) * w[:, None]  # [m, C]
    return er[:, None, :, None] * ec[None, :, None, :] / (n * n)


def correlation_functional_unshifted(n: int, crop: int) -> np.ndarray:
    """The same lag map written against UNshifted packed operands.

    Derivation: fftshift rolls rows by n/2, so position r holds original row
    ro = r - n/2 and e^{2 pi i r ly/n} = (-1)^ly e^{2 pi i ro ly/n}; it rolls the
    n/2+1 columns by s = (n/2+1)//2, so original column co sits at position
    c = (co + s) mod (n/2+1), which fixes both its phase and its C2R weight.
    """
    m = n // 2 + 1
    s = m // 2
    lag = crop_indices(n, crop)
    pos = (np.arange(m) + s) % m
    w = np.where((pos == 0) | (pos == n // 2), 1.0, 2.0)
    sign = np.where(lag % 2 == 0, 1.0, -1.0)
    er = np.exp(2j * np.pi * np.outer(np.arange(n), lag) / n) * sign[None, :]
    ec = np.exp(2j * np.pi * np.outer(pos, lag) / n) * w[:, None]
    return er[:, None, :, None] * ec[None, :, None, :] / (n * n)


def correlate_direct(templates_pre, particles_pre, crop: int):
    """O(n^2 C^2) direct evaluation of the same operator via the explicit functional."""
    t = np.asarray(templates_pre, np.complex128)
    p = np.asarray(particles_pre, np.complex128)
    n = t.shape[-2]
    phi = correlation_functional(n, crop)
    A, B = t.shape[0], p.shape[0]
    out = np.empty((A, B, crop, crop))
    for a in range(A):
        for b in range(B):
            q = t[a] * p[b]
            out[a, b] = np.einsum("rc,rcij->ij", q, phi).real
    return out
