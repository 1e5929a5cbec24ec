"""Host-side setup of the search: masks, FSC weights, direction matrices,
local orientation map, psi sampling grids and static CTF tables.

Everything here is small, computed once per dataset or per call on the host
in NumPy with the reference's float32/float64 conventions, and handed to the
HIP kernels as plain arrays.
"""
from __future__ import annotations

import math
from typing import List, Sequence, Tuple

import numpy as np

PI_V606 = 3.14159265359  # literal pi of v606:170-172, 345


def euler_angles_to_matrix(alpha: float, beta: float, gamma: float) -> List[List[float]]:
    """RELION ZYZ Euler matrix in degrees (v606:168-187)."""
    a = alpha / 180.0 * PI_V606
    b = beta / 180.0 * PI_V606
    g = gamma / 180.0 * PI_V606
    ca, cb, cg, sa, sb, sg = math.cos(a), math.cos(b), math.cos(g), math.sin(a), math.sin(b), math.sin(g)
    cc, cs, sc, ss = cb * ca, cb * sa, sb * ca, sb * sa
    return [[cg * cc - sg * sa, cg * cs + sg * ca, -cg * sb],
            [-sg * cc - cg * sa, -sg * cs + cg * ca, sg * sb],
            [sc, ss, cb]]


def euler_batch(rot, tilt) -> np.ndarray:
    """[K, 3, 3] float64 Euler(rot, tilt, 0) with the same literal pi."""
    a = np.asarray(rot, np.float64) / 180.0 * PI_V606
    b = np.asarray(tilt, np.float64) / 180.0 * PI_V606
    ca, cb, sa, sb = np.cos(a), np.cos(b), np.sin(a), np.sin(b)
    z = np.zeros_like(a)
    return np.stack([np.stack([cb * ca, cb * sa, -sb], -1),
                     np.stack([-sa, ca, z], -1),
                     np.stack([sb * ca, sb * sa, cb], -1)], -2)


def rotation_matrices_t(rot, tilt) -> np.ndarray:
    """[M, 3, 3] float32 Euler(rot, tilt, 0)^T as read_model_angles stores them (v606:214-215)."""
    return np.stack([np.asarray(euler_angles_to_matrix(r, t, 0.0), np.float32).T
                     for r, t in zip(rot, tilt)]).astype(np.float32)


def calculate_angular_distance(rot1, tilt1, psi1, rot2, tilt2, psi2) -> float:
    """Mean of the three row-vector angles between two Euler matrices (v606:331-348)."""
    e1 = euler_angles_to_matrix(rot1, tilt1, psi1)
    e2 = euler_angles_to_matrix(rot2, tilt2, psi2)
    total = 0.0
    for i in range(3):
        d = sum(float(a) * float(b) for a, b in zip(e1[i], e2[i]))
        total += math.acos(min(max(d, -1.0), 1.0)) * 180.0 / PI_V606
    return total / 3.0


def local_orientation_map(prior_rot, prior_tilt, model_rot, model_tilt, local_range) -> List[List[int]]:
    """build_local_orientation_entries (v606:351-378), vectorised.

    A direction is kept when |d rot| <= r and |d tilt| <= r (wrapped) and the
    mean row-vector angle with psi = 0 is <= r.  Indices ascend.
    """
    mr = np.asarray(model_rot, np.float64)
    mt = np.asarray(model_tilt, np.float64)
    em = euler_batch(mr, mt)
    ep = euler_batch(prior_rot, prior_tilt)
    out: List[List[int]] = []
    for i in range(len(ep)):
        rd = np.fabs((mr - float(prior_rot[i]) + 180.0) % 360.0 - 180.0)
        td = np.fabs((mt - float(prior_tilt[i]) + 180.0) % 360.0 - 180.0)
        cand = np.nonzero((rd <= local_range) & (td <= local_range))[0]
        if len(cand) == 0:
            out.append([])
            continue
        dots = np.einsum("kij,ij->ki", em[cand], ep[i])
        ang = np.arccos(np.clip(dots, -1.0, 1.0)) * 180.0 / PI_V606
        dist = (ang[:, 0] + ang[:, 1] + ang[:, 2]) / 3.0
        out.append([int(k) for k in cand[np.abs(dist) <= local_range]])
    return out


def generate_soft_mask(shape: Tuple[int, int], radius: int, edge_width: int) -> np.ndarray:
    """Cosine-edged disc centred at shape//2 (v606:446-457)."""
    center = np.asarray(shape) // 2
    y, x = np.indices(shape)
    d = np.sqrt((x - center[1]) ** 2 + (y - center[0]) ** 2)
    mask = np.zeros_like(d)
    mask[d <= radius] = 1.0
    if edge_width > 0:
        edge = (d > radius) & (d <= radius + edge_width)
        mask[edge] = 0.5 * (1.0 + np.cos(np.pi * (d[edge] - radius) / edge_width))
    return mask.astype(np.float32)


def expand_radial_values(values, size: int) -> np.ndarray:
    """Radial table -> centred [size, size] map by truncated radius (v606:460-468)."""
    v = np.asarray(values, np.float32).reshape(-1)
    c = size // 2
    y, x = np.indices((size, size))
    r = np.sqrt((y - c) ** 2 + (x - c) ** 2).astype(np.int64)
    out = np.zeros((size, size), np.float32)
    ok = r < v.shape[0]
    out[ok] = v[r[ok]]
    return out


def centered_to_packed(centered: np.ndarray) -> np.ndarray:
    """fftshift-centred [n, n] plane -> packed rfft2 layout [n, n/2+1] (v606:505-533)."""
    n = centered.shape[-1]
    h = n // 2
    rows = np.concatenate((np.arange(h, n), np.arange(0, h)))
    cols = np.concatenate((np.arange(h, n), np.zeros(1, np.int64)))
    return np.take(np.take(centered, rows, axis=-2), cols, axis=-1)


def packed_fsc(fsc_values, new_box: int) -> np.ndarray:
    """Packed FSC weights (read_fsc_map + centered_full_to_packed, v606:471-491, 1644-1650)."""
    return np.ascontiguousarray(centered_to_packed(expand_radial_values(fsc_values, new_box)))


def read_fsc_values(fsc_filename: str, original_box: int, ignore_fsc: bool) -> np.ndarray:
    """Second column of an FSC text file, or ones (v606:471-491)."""
    if not fsc_filename:
        return np.ones(original_box, np.float32)
    parsed = []
    with open(fsc_filename, "r", encoding="utf-8", errors="replace") as fh:
        for line in fh:
            tok = line.split()
            if tok:
                parsed.append(1.0 if ignore_fsc else float(tok[1]))
    if not parsed:
        raise ValueError(f"No FSC values found in {fsc_filename}")
    return np.asarray(parsed, np.float32)


def psi_sampling_for_record(particle_psi, psi_step, do_local_search, local_range):
    """Psi angles searched for one particle (v606:1163-1190).

    Returns (fake, effective); effective maps |angle| < 0.5 to 0.5.
    """
    num_psi = int(360 // psi_step)
    if num_psi < 1:
        num_psi = 1
    tmp = 999999.0
    if do_local_search:
        tmp = local_range // psi_step
        local_count = int(tmp * 2 + 1)
        if local_count < num_psi:
            num_psi = local_count
    nearest = round(((particle_psi + 360.0) % 360.0) / psi_step) * psi_step
    nearest %= 360.0
    fake, eff = [], []
    for i in range(int(num_psi)):
        a = psi_step * i
        if do_local_search:
            a = nearest - tmp * psi_step + i * psi_step
        fake.append(float(a))
        eff.append(float(0.5 if abs(a) < 0.5 else a))
    return fake, eff


def psi_grid_params(size: int, effective_angles: Sequence[float]) -> np.ndarray:
    """[P, 6] float32 (c, -s, s, c, off_y, off_x) of PsiGridBank (v606:1193-1239)."""
    center = np.asarray((size, size), np.float32) / np.float32(2.0)
    out = np.empty((len(effective_angles), 6), np.float32)
    for i, a in enumerate(effective_angles):
        theta = np.deg2rad(float(np.round(a, 6)))
        c = np.float32(np.cos(theta))
        s = np.float32(np.sin(theta))
        m = np.asarray([[c, -s], [s, c]], np.float32)
        off = (center - m @ center).astype(np.float32)
        out[i] = (c, -s, s, c, off[0], off[1])
    return out


class CTFTables:
    """Static float32 terms of PackedCTFTorch (v606:695-776): x, y, u4, ncurve
    on the packed [n, n/2+1] layout and the scalars k1, k2, k3, k5, kk."""

    def __init__(self, size, pixel_size, cs, voltage, kk, amp_const=0.1, bfactor=0.0, phase_shift=0.0):
        if not 0.0 <= amp_const <= 1.0:
            raise ValueError("Amplitude contrast must be in [0, 1]")
        if bfactor != 0.0:
            raise ValueError("B-factor != 0 is not used by the search path")
        f32 = np.float32
        n = int(size)
        self.size = n
        local_cs = f32(cs) * f32(1e7)
        local_kv = f32(voltage) * f32(1e3)
        lam = f32(12.2643247) / np.sqrt(local_kv * (f32(1.0) + local_kv * f32(0.978466e-6)))
        self.k1 = float(f32(np.pi) * lam)
        self.k2 = float(f32(np.pi / 2.0) * local_cs * lam ** f32(3.0))
        self.k3 = float(np.arctan(f32(amp_const) / np.sqrt(f32(1.0) - f32(amp_const) ** 2)).astype(f32))
        self.k5 = float(np.radians(f32(phase_shift)).astype(f32))
        self.kk = float(kk)
        px = f32(pixel_size)
        cf = (np.arange(n, dtype=f32) - f32(n) / f32(2.0)) / (f32(n) * px)
        yf = np.concatenate((cf[n // 2:], cf[:n // 2]))
        xf = np.concatenate((cf[n // 2:], cf[:1]))
        x, y = np.meshgrid(xf, yf, indexing="xy")
        x, y = x.astype(f32), y.astype(f32)
        u2 = x * x + y * y
        u4 = u2 * u2
        ss = np.sqrt(u2).astype(f32)
        signal = np.exp(f32(-78.7757) * ss * ss + f32(-12.9121) * ss + f32(1.28732), dtype=f32) / (f32(kk) + f32(1.0))
        ncurve = (np.exp(f32(-9.32) * ss * ss + f32(2.65) * ss + f32(0.01908), dtype=f32) / signal).astype(f32)
        self.table = np.ascontiguousarray(np.stack([x, y, u4, ncurve]).astype(f32))  # [4, n, n/2+1]

    @property
    def kconst(self) -> np.ndarray:
        return np.asarray([self.k1, self.k2, self.k3, self.k5, self.kk], np.float32)


def crop_functional(n: int, crop: int):
    """The v606 crop as a separable linear functional of the UNshifted packed
    product Q = T * P (correlate_templates_and_particle, v606:1418-1449 with
    the pre-shift of v606:535-537 folded in):

        score[y][x] = 1024 * Re( (W^T Q E)[y][x] ),
        W[r][y] = (-1)^l_y exp(2 pi i r l_y / n),
        E[c][x] = w(pos_c) exp(2 pi i pos_c l_x / n),  pos_c = (c + (n/2+1)//2) mod (n/2+1),

    w = 1 for pos in {0, n/2} (C2R drops the imaginary part of those
    columns) else 2, l the crop lags [n-C/2 .. n-1, 0 .. C/2-1].  Returns
    complex128 W [n, C] and E [n/2+1, C]."""
    n, crop = int(n), int(crop)
    m = n // 2 + 1
    h = crop // 2
    lags = np.concatenate((np.arange(n - h, n), np.arange(0, h)))
    pos = (np.arange(m) + m // 2) % m
    w = np.where((pos == 0) | (pos == n // 2), 1.0, 2.0)
    sign = np.where(lags % 2 == 0, 1.0, -1.0)
    W = np.exp(2j * np.pi * (np.outer(np.arange(n), lags) % n) / n) * sign[None, :]
    E = np.exp(2j * np.pi * (np.outer(pos, lags) % n) / n) * w[:, None]
    return W, E


class SmallShiftTables:
    """Coefficient set and split lag table of the small-shift precision search.

    The v606 score at integer lag (ly, lx) is a fixed linear functional of
    the unshifted packed product Q = T * P (correlate_templates_and_particle,
    v606:1418-1449 followed by the C2R of torch.fft.irfft2): with
    pos = (c + (n/2+1)//2) mod (n/2+1) the column after the pre-shift roll,
    w = 1 for pos in {0, n/2} and 2 otherwise (C2R drops the imaginary part
    of those columns) and (-1)^ly from the row roll by n/2,

        score(ly, lx) = 1024 * sum_{r,c} Re(Q[r][c] * w * (-1)^ly * exp(2 pi i (r ly + pos lx) / n)).

    Only coefficients with |k| <= radius (frequency fy = r or r - n, fx = c,
    with the packed Nyquist column at fx = -n/2) are kept and the lags run
    over [-half, half]^2, row-major.  The real GEMM operand B[2k][s] =
    Re Phi, B[2k+1][s] = -Im Phi is split into 16-bit hi + lo (``operand``
    "bf16" or "fp16", both round to nearest even; |B| <= 2 needs no scaling
    in either) and stored in the v_mfma_f32_16x16x32_{bf16,f16} B-operand
    order [lag tile][k-step][lane][8]: lane l of k-step ks holds rows
    32 ks + 8 (l / 16) + 0..7 of column 16 tile + l % 16.
    """

    LAG_TILES = 4

    OPERANDS = {"bf16": 0, "fp16": 1}

    def __init__(self, n: int, radius: float, half: int = 3, operand: str = "bf16"):
        n = int(n)
        half = int(half)
        if operand not in self.OPERANDS:
            raise ValueError(f"operand {operand!r} not in {sorted(self.OPERANDS)}")
        self.operand = operand
        side = 2 * half + 1
        if side * side > 16 * self.LAG_TILES:
            raise ValueError(f"half {half}: at most {16 * self.LAG_TILES} lags")
        m = n // 2 + 1
        self.n, self.half, self.radius = n, half, float(radius)
        r = np.arange(n)
        fy = np.where(r < n // 2, r, r - n)
        c = np.arange(m)
        fx = np.where(c < n // 2, c, -(n // 2))
        inside = (fy[:, None] ** 2 + fx[None, :] ** 2) <= self.radius ** 2
        self.kidx = np.flatnonzero(inside).astype(np.int32)
        self.K = int(self.kidx.size)
        self.Kpad = max(16, -(-self.K // 16) * 16)
        # inverse map: slot of each packed entry (-1 = not kept), [n][n/2+1]
        self.kslot = np.full(n * m, -1, np.int32)
        self.kslot[self.kidx] = np.arange(self.K, dtype=np.int32)
        self.kslot = self.kslot.reshape(n, m)
        rows, cols = self.kidx // m, self.kidx % m
        pos = (cols + m // 2) % m
        wt = np.where((pos == 0) | (pos == n // 2), 1.0, 2.0)
        lags = np.arange(-half, half + 1)
        ly = np.repeat(lags, side)
        lx = np.tile(lags, side)
        sgn = 1.0 - 2.0 * (np.abs(ly) % 2)
        phase = (2.0 * np.pi / n) * (np.outer(rows, ly) + np.outer(pos, lx))
        phi = (wt[:, None] * sgn[None, :]) * np.exp(1j * phase)          # [K, S]
        S = side * side
        breal = np.zeros((2 * self.Kpad, 16 * self.LAG_TILES), np.float64)
        breal[0:2 * self.K:2, :S] = phi.real
        breal[1:2 * self.K:2, :S] = -phi.imag
        self.phi = phi
        # last lag as float32 (Re Phi, -Im Phi) pairs: 49 = 3 x 16 + 1 lags run 3
        # tiles on the MFMAs and this one on the VALU
        self.phi_last = np.zeros((self.Kpad, 2), np.float32)
        self.phi_last[:self.K, 0] = phi[:, S - 1].real
        self.phi_last[:self.K, 1] = -phi[:, S - 1].imag
        hi, lo = _split_bf16(breal) if operand == "bf16" else _split_f16(breal)
        self.b_hi = _mfma_b_layout(hi, self.Kpad, self.LAG_TILES)
        self.b_lo = _mfma_b_layout(lo, self.Kpad, self.LAG_TILES)

    @property
    def n_lags(self) -> int:
        return (2 * self.half + 1) ** 2


def _bf16_bits(x32: np.ndarray) -> np.ndarray:
    """float32 -> bfloat16 bit patterns, round to nearest even (finite inputs)."""
    u = np.ascontiguousarray(x32, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def _bf16_value(bits: np.ndarray) -> np.ndarray:
    return (bits.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _split_bf16(x: np.ndarray):
    hi = _bf16_bits(x.astype(np.float32))
    lo = _bf16_bits((x - _bf16_value(hi)).astype(np.float32))
    return hi, lo


def _split_f16(x: np.ndarray):
    """x = hi + lo in IEEE half, each rounded to nearest even (bit patterns)."""
    hi = x.astype(np.float16)
    lo = (x - hi.astype(np.float64)).astype(np.float16)
    return hi.view(np.uint16), lo.view(np.uint16)


def _mfma_b_layout(bits: np.ndarray, kpad: int, tiles: int) -> np.ndarray:
    """[2 Kpad, 16 tiles] -> [tiles, Kpad/16, 64, 8] (lane l: rows 8 (l/16) + e, column l % 16)."""
    ks = kpad // 16
    b = bits.reshape(ks, 4, 8, tiles, 16)          # [ks][g][e][tile][i]
    b = b.transpose(3, 0, 1, 4, 2)                 # [tile][ks][g][i][e]
    return np.ascontiguousarray(b.reshape(tiles, ks, 64, 8))

