"""Oracle: atom -> voxel radial-exponential splat and the rigid-body transform.

Test infrastructure only (see ``oracle/__init__.py``).
"""
from __future__ import annotations

import math

import numpy as np

# pdb2mrc_gpu_ver_fp32_v3.py:13-16 -- atomic number per element symbol.
ELEMENT_Z = {"H": 1.0, "C": 6.0, "N": 7.0, "O": 8.0, "P": 15.0, "S": 16.0}
ELEMENT_MASS = {"H": 1.00794, "C": 12.0107, "N": 14.00674, "O": 15.9994,
                "P": 30.973761, "S": 32.066}


def splat_parameters(apix: float, res: float):
    """width, rp, kn exactly as pdb2mrc_gpu_ver_fp32_v3.py:53-58."""
    width = int(round(res * 3.0 / apix))
    rp = float(np.power(np.pi / (res / apix), 2))
    kn = float(np.power(rp / np.pi, 1.5))
    return width, rp, kn


def _round_half_away(v: np.ndarray) -> np.ndarray:
    """pdb2mrc_gpu_ver_fp32_v3.py:35-38: sign(v)*floor(|v|+0.5) in float64."""
    v = v.astype(np.float64)
    return (np.sign(v) * np.floor(np.abs(v) + 0.5)).astype(np.int64)


def density_from_coords(coordinates, atomic_numbers, boxsize, apix, res,
                        accumulate=np.float32):
    """Restatement of pdb2mrc_gpu_ver_fp32_v3.py:41-98 (returns [z, y, x]).

    ``accumulate`` selects the voxel accumulator type: float32 mirrors the
    reference's ``index_add_``; float64 gives the exact sum of the same
    float32 per-atom values.
    """
    boxsize = int(boxsize)
    apix, res = float(apix), float(res)
    if boxsize < 1 or not np.isfinite(apix) or not np.isfinite(res) or apix <= 0 or res <= 0:
        raise ValueError("boxsize, apix and resolution must be positive and finite")
    width, rp, kn = splat_parameters(apix, res)
    if width < 3:
        raise ValueError("insufficient sampling")
    xyz = np.asarray(coordinates, dtype=np.float32).reshape(-1, 3)
    numbers = np.asarray(atomic_numbers, dtype=np.float32).reshape(-1)
    if len(xyz) != len(numbers):
        raise ValueError("Coordinate and atom count differ")
    vol = np.zeros(boxsize ** 3, dtype=accumulate)
    if len(xyz) == 0:
        return vol.reshape(boxsize, boxsize, boxsize).astype(np.float32)
    if not np.isfinite(xyz).all():
        raise ValueError("PDB coordinates must be finite")
    pos = (xyz.astype(np.float64) / apix + boxsize // 2).astype(np.float32)
    lower = _round_half_away(pos - np.float32(width))
    upper = _round_half_away(pos + np.float32(width))
    valid = (lower < boxsize).all(1) & (upper >= 0).all(1) & (numbers != 0)
    pos, numbers = pos[valid], numbers[valid]
    lower = np.maximum(lower[valid], 0)
    upper = np.minimum(upper[valid], boxsize - 1)
    axis = np.arange(2 * width + 1)
    zz, yy, xx = np.meshgrid(axis, axis, axis, indexing="ij")
    offsets = np.stack((xx.ravel(), yy.ravel(), zz.ravel()), 1)  # (K, 3) as (x, y, z)
    chunk = max(1, 200000 // len(offsets))
    for s in range(0, len(pos), chunk):
        e = s + chunk
        vox = lower[s:e, None, :] + offsets[None]
        ok = (vox < upper[s:e, None, :]).all(2)
        d = vox.astype(np.float32) - pos[s:e, None, :]
        sq = d * d
        r = np.sqrt((sq[..., 0] + sq[..., 1]) + sq[..., 2])
        expo = (-r.astype(np.float64) * rp).astype(np.float32)
        val = ((numbers[s:e, None].astype(np.float64) * kn)
               * np.exp(expo).astype(np.float64)).astype(np.float32)
        idx = vox[..., 0] + boxsize * (vox[..., 1] + boxsize * vox[..., 2])
        np.add.at(vol, idx[ok], val[ok].astype(accumulate))
    return vol.reshape(boxsize, boxsize, boxsize).astype(np.float32)


def density_brute_force(coordinates, atomic_numbers, boxsize, apix, res):
    """Per-voxel double loop of the same rule; for tiny boxes only (pinning)."""
    width, rp, kn = splat_parameters(apix, res)
    vol = np.zeros((boxsize, boxsize, boxsize), np.float64)
    for (x, y, z), zn in zip(np.asarray(coordinates, np.float32), atomic_numbers):
        if zn == 0:
            continue
        p = [np.float32(float(c) / apix + boxsize // 2) for c in (x, y, z)]
        lo, hi = [], []
        for c in p:
            lv, uv = float(np.float32(c - np.float32(width))), float(np.float32(c + np.float32(width)))
            lo.append(int(math.copysign(math.floor(abs(lv) + 0.5), lv)))
            hi.append(int(math.copysign(math.floor(abs(uv) + 0.5), uv)))
        if any(l >= boxsize for l in lo) or any(h < 0 for h in hi):
            continue
        lo = [max(l, 0) for l in lo]
        hi = [min(h, boxsize - 1) for h in hi]
        for vz in range(lo[2], hi[2]):
            for vy in range(lo[1], hi[1]):
                for vx in range(lo[0], hi[0]):
                    r = math.sqrt((vx - p[0]) ** 2 + (vy - p[1]) ** 2 + (vz - p[2]) ** 2)
                    vol[vz, vy, vx] += zn * kn * math.exp(-r * rp)
    return vol


def rotvec_matrix(rotvec_deg) -> np.ndarray:
    """Rotation matrix of a rotation vector in degrees (Rodrigues' formula).

    Equals scipy ``Rotation.from_rotvec(deg2rad(v)).as_matrix()`` used at
    pdb_text.py:140.
    """
    v = np.deg2rad(np.asarray(rotvec_deg, dtype=np.float64))
    theta = float(np.linalg.norm(v))
    if theta == 0.0:
        return np.eye(3)
    k = v / theta
    kx = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) * math.cos(theta) + (1 - math.cos(theta)) * np.outer(k, k) + math.sin(theta) * kx


def transform_coordinates(coords, pose, pivot=None) -> np.ndarray:
    """pdb_text.py:123-142: (X - pivot) @ R + pivot + t, pivot = mean by default."""
    coords = np.asarray(coords)
    pose = np.asarray(pose, dtype=np.float64)
    center = np.mean(coords, axis=0) if pivot is None else np.asarray(pivot, np.float64)
    return (coords - center) @ rotvec_matrix(pose[:3]) + center + pose[3:]


def quantize_8_3(xyz) -> np.ndarray:
    """The PDB 8.3 text round trip: float(f"{x:8.3f}") per coordinate."""
    return np.array([[float(f"{float(v):8.3f}") for v in row] for row in np.asarray(xyz)],
                    dtype=np.float32)
