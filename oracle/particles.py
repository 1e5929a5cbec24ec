"""Oracle: packed CTF, psi sampling, in-plane rotation and particle preparation.

Test infrastructure only (see ``oracle/__init__.py``).
"""
from __future__ import annotations

import numpy as np

from .fourier import centered_full_to_packed, pre_shift


class PackedCTF:
    """Restatement of PackedCTFNumPy (v606:549-692), float32 throughout."""

    def __init__(self, size, pixel_size, cs, voltage, kk, amp_const=0.1):
        f32 = np.float32
        self.size = n = int(size)
        self.kk = f32(kk)
        local_cs = f32(f32(cs) * f32(1e7))
        local_kv = f32(f32(voltage) * f32(1e3))
        lam = f32(12.2643247) / np.sqrt(local_kv * (f32(1.0) + local_kv * f32(0.978466e-6)))
        self.k1 = f32(np.pi) * lam
        self.k2 = f32(np.pi / 2.0) * local_cs * lam ** f32(3.0)
        ac = f32(amp_const)
        self.k3 = np.arctan(ac / np.sqrt(f32(1.0) - ac * ac)).astype(f32)
        self.k5 = f32(0.0)
        px = f32(pixel_size)
        cf = (np.arange(n, dtype=f32) - f32(n) / f32(2.0)) / (f32(n) * px)
        yf = np.concatenate((cf[n // 2:], cf[:n // 2]))
        xf = np.concatenate((cf[n // 2:], cf[:1]))
        x, y = np.meshgrid(xf, yf, indexing="xy")
        self.x, self.y = x.astype(f32), y.astype(f32)
        u2 = self.x * self.x + self.y * self.y
        self.u4 = u2 * u2
        ss = np.sqrt(u2).astype(f32)
        signal = np.exp(f32(-78.7757) * ss * ss + f32(-12.9121) * ss + f32(1.28732),
                        dtype=f32) / (self.kk + f32(1.0))
        self.ncurve = (np.exp(f32(-9.32) * ss * ss + f32(2.65) * ss + f32(0.01908), dtype=f32)
                       / signal).astype(f32)

    def one(self, dfu, dfv, dfa):
        """v606:646-682 (bfactor = 0, scale = 1)."""
        f32 = np.float32
        dfu, dfv = f32(dfu), f32(dfv)
        ang = np.radians(f32(dfa)).astype(f32)
        s, c = np.sin(ang).astype(f32), np.cos(ang).astype(f32)
        axx = -dfu * c * c - dfv * s * s
        ayy = -dfu * s * s - dfv * c * c
        axy = -dfu * s * c + dfv * s * c
        gamma = (self.k1 * (axx * self.x * self.x + f32(2.0) * axy * self.x * self.y
                            + ayy * self.y * self.y) + self.k2 * self.u4 - self.k5 - self.k3)
        v = -np.sin(gamma, dtype=f32)
        v = np.where(np.abs(v) < f32(1e-8), np.sign(v) * f32(1e-8), v).astype(f32)
        return (v * np.sqrt(f32(1.0) / (self.ncurve + self.kk * v * v), dtype=f32)).astype(f32)


def psi_sampling_for_record(particle_psi, psi_step, do_local_search, local_range):
    """v606:1163-1190, including the |angle| < 0.5 -> 0.5 quirk."""
    num_psi = int(360 // psi_step)
    if num_psi < 1:
        num_psi = 1
    tmp = 999999.0
    if do_local_search:
        tmp = local_range // psi_step
        local_count = int(tmp * 2 + 1)
        if local_count < num_psi:
            num_psi = local_count
    fake, eff = [], []
    nearest = round(((particle_psi + 360.0) % 360.0) / psi_step) * psi_step
    nearest %= 360.0
    for i in range(int(num_psi)):
        a = psi_step * i
        if do_local_search:
            a = nearest - tmp * psi_step + i * psi_step
        fake.append(float(a))
        eff.append(float(0.5 if abs(a) < 0.5 else a))
    return fake, eff


def rotation_source(size, angle):
    """v606:1213-1239: source (y, x) for every output pixel, float32 c/s/offset."""
    theta = np.deg2rad(float(np.round(angle, 6)))
    c, s = np.float32(np.cos(theta)), np.float32(np.sin(theta))
    m = np.asarray([[c, -s], [s, c]], np.float32)
    center = np.asarray((size, size), np.float32) / np.float32(2.0)
    off = (center - m @ center).astype(np.float32)
    yy, xx = np.indices((size, size)).astype(np.float64)
    sy = float(c) * yy + float(-s) * xx + float(off[0])
    sx = float(s) * yy + float(c) * xx + float(off[1])
    return sy, sx


def bilinear_zeros(img, sy, sx):
    """F.grid_sample semantics: bilinear, padding_mode='zeros', align_corners=True."""
    H, W = img.shape
    y0 = np.floor(sy).astype(np.int64)
    x0 = np.floor(sx).astype(np.int64)
    ty, tx = sy - y0, sx - x0
    out = np.zeros(sy.shape, np.float64)
    for dy in (0, 1):
        wy = ty if dy else 1 - ty
        yi = y0 + dy
        for dx in (0, 1):
            wx = tx if dx else 1 - tx
            xi = x0 + dx
            ok = (yi >= 0) & (yi < H) & (xi >= 0) & (xi < W)
            v = img[np.clip(yi, 0, H - 1), np.clip(xi, 0, W - 1)]
            out += np.where(ok, wy * wx * v, 0.0)
    return out


def prepare_particle(img, effective_angles, new_box, mask, packed_ctf, dtype=np.float32):
    """ParticleFourierPreparer.prepare (v606:1340-1374) -> pre-shifted [P, n, n/2+1]."""
    N = img.shape[0]
    out = []
    for a in effective_angles:
        sy, sx = rotation_source(N, a)
        r = bilinear_zeros(np.asarray(img, np.float64), sy, sx).astype(dtype)
        if mask is not None:
            r = r * mask.astype(dtype)
        F = np.fft.fftshift(np.fft.fft2(r))
        if new_box != N:
            lo, hi = N // 2 - new_box // 2, N // 2 + new_box // 2
            F = F[lo:hi, lo:hi]
        p = np.conj(centered_full_to_packed(F) * packed_ctf.astype(dtype))
        p = p / np.linalg.norm(p.ravel())
        out.append(pre_shift(p))
    cdt = np.complex64 if dtype == np.float32 else np.complex128
    return np.stack(out).astype(cdt)
