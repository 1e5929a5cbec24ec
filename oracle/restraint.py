"""Oracle: the geometric restraint (func_check_boundary_for_testing_v8.py:33-126).

Test infrastructure only (see ``oracle/__init__.py``).  The distance
transform is brute force (every voxel against every mask voxel), for tiny
grids; scipy.ndimage.distance_transform_edt pins it in the tests.
"""
from __future__ import annotations

import numpy as np

from .splat import density_from_coords


def edt_brute_force(mask: np.ndarray) -> np.ndarray:
    """Euclidean distance (voxels) from every voxel to the nearest True voxel
    of a 3-D boolean grid (distance_transform_edt(~mask))."""
    pts = np.argwhere(mask).astype(np.float64)
    grid = np.indices(mask.shape).reshape(3, -1).T.astype(np.float64)
    d2 = ((grid[:, None, :] - pts[None, :, :]) ** 2).sum(-1).min(1)
    return np.sqrt(d2).reshape(mask.shape)


def rotvec_matrix_f32(rotvec_deg) -> np.ndarray:
    """scipy Rotation.from_rotvec(deg2rad(float32 angles)).as_matrix() as float32 (:42-43)."""
    v = np.deg2rad(np.asarray(rotvec_deg, np.float32)).astype(np.float64)
    t = float(np.linalg.norm(v))
    if t == 0.0:
        return np.eye(3, dtype=np.float32)
    k = v / t
    kx = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    m = np.eye(3) * np.cos(t) + (1 - np.cos(t)) * np.outer(k, k) + np.sin(t) * kx
    return m.astype(np.float32)


def rotate_coords(coords, pose, pivot=None) -> np.ndarray:
    """rotate_coords_rotvec (:33-45): float32 row vectors about the mean."""
    xyz = np.asarray(coords, np.float32)
    c = np.mean(xyz, axis=0) if pivot is None else np.asarray(pivot, np.float32)
    return ((xyz - c) @ rotvec_matrix_f32(pose[:3]) + c + np.asarray(pose[3:], np.float32)).astype(np.float32)


def restraint(chain_xyz, chain_z, rest_xyz, rest_z, pose, box, apix, threshold=1.0, pivot=None):
    """GeometryRestraint(...).evaluate(pose) (:62-126): (overlap, distance A)."""
    body = density_from_coords(rest_xyz, rest_z, box, apix, 2.0 * apix, accumulate=np.float64) > threshold
    if not body.any():
        return 0, 0.0
    moved = rotate_coords(chain_xyz, pose, pivot)
    chain = density_from_coords(moved, chain_z, box, apix, 2.0 * apix, accumulate=np.float64) > threshold
    overlap = int(np.count_nonzero(chain & body))
    if overlap:
        return overlap, 0.0
    if chain.any():
        return 0, float(edt_brute_force(body)[chain].min()) * apix
    zz, yy, xx = np.nonzero(body)
    return 0, float(np.sqrt(np.min((zz + 1) ** 2 + yy ** 2 + xx ** 2))) * apix
