"""Pins the oracle's splat (pdb2mrc_gpu_ver_fp32_v3.py:41-98) and rigid transform
(pdb_text.py:123-142) against closed forms and brute force."""
import math

import numpy as np
import pytest

from oracle import splat as osp


def test_parameters_closed_form():
    w, rp, kn = osp.splat_parameters(1.5, 3.0)
    assert w == 6
    assert rp == pytest.approx((math.pi / 2.0) ** 2)
    assert kn == pytest.approx((rp / math.pi) ** 1.5)


def test_single_atom_closed_form():
    box, apix, res = 24, 1.0, 2.0
    w, rp, kn = osp.splat_parameters(apix, res)
    xyz = np.array([[0.25, -0.5, 1.0]], np.float32)  # A, relative to the box centre
    vol = osp.density_from_coords(xyz, [8.0], box, apix, res, accumulate=np.float64)
    p = xyz[0].astype(np.float64) / apix + box // 2
    lo = [int(math.copysign(math.floor(abs(c - w) + 0.5), c - w)) for c in p]
    hi = [int(math.copysign(math.floor(abs(c + w) + 0.5), c + w)) for c in p]
    want = np.zeros_like(vol, dtype=np.float64)
    for z in range(lo[2], hi[2]):
        for y in range(lo[1], hi[1]):
            for x in range(lo[0], hi[0]):
                r = math.sqrt((x - p[0]) ** 2 + (y - p[1]) ** 2 + (z - p[2]) ** 2)
                want[z, y, x] = 8.0 * kn * math.exp(-rp * r)   # exp of r, not r^2
    np.testing.assert_allclose(vol, want, rtol=2e-6, atol=1e-9)


def test_vectorised_equals_brute_force():
    rng = np.random.default_rng(0)
    box, apix, res = 12, 1.0, 3.0
    xyz = rng.uniform(-8, 8, (10, 3)).astype(np.float32)
    z = rng.choice([0.0, 1.0, 6.0, 16.0], 10)
    a = osp.density_from_coords(xyz, z, box, apix, res, accumulate=np.float64)
    b = osp.density_brute_force(xyz, z, box, apix, res)
    np.testing.assert_allclose(a, b, rtol=2e-6, atol=1e-9)


def test_bounds_and_skips():
    box, apix, res = 16, 1.0, 2.0
    vol = osp.density_from_coords([[7.4, 7.4, 7.4]], [6.0], box, apix, res)
    # upper bound is exclusive and clipped at box-1: the last plane is never written
    assert vol[-1].sum() == 0 and vol[:, -1].sum() == 0 and vol[:, :, -1].sum() == 0
    assert vol.sum() > 0
    assert osp.density_from_coords([[0, 0, 0]], [0.0], box, apix, res).sum() == 0  # Z = 0
    assert osp.density_from_coords([[500, 0, 0]], [6.0], box, apix, res).sum() == 0  # outside
    with pytest.raises(ValueError):
        osp.density_from_coords([[0, 0, 0]], [6.0], box, 1.0, 0.5)  # width < 3
    with pytest.raises(ValueError):
        osp.density_from_coords([[np.nan, 0, 0]], [6.0], box, apix, res)
    with pytest.raises(ValueError):
        osp.density_from_coords([[0, 0, 0]], [6.0, 7.0], box, apix, res)


def test_rotvec_matches_closed_form_rotations():
    # 90 degrees about z maps x -> y (column-vector convention of from_rotvec)
    R = osp.rotvec_matrix([0, 0, 90])
    np.testing.assert_allclose(R @ [1, 0, 0], [0, 1, 0], atol=1e-12)
    # rotation vector of length theta about axis k: R k = k, R orthonormal, det 1
    k = np.array([1.0, 2.0, -0.5])
    k /= np.linalg.norm(k)
    R = osp.rotvec_matrix(k * 37.0)
    np.testing.assert_allclose(R @ k, k, atol=1e-12)
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-12)
    assert np.linalg.det(R) == pytest.approx(1.0)
    assert math.degrees(math.acos((np.trace(R) - 1) / 2)) == pytest.approx(37.0)


def test_transform_row_convention_and_pivot():
    X = np.array([[1.0, 0, 0], [-1.0, 0, 0], [0, 2.0, 0]])
    out = osp.transform_coordinates(X, [0, 0, 90, 1, 2, 3])
    c = X.mean(0)
    R = osp.rotvec_matrix([0, 0, 90])
    np.testing.assert_allclose(out, (X - c) @ R + c + [1, 2, 3])
    # zero pose is the identity
    np.testing.assert_allclose(osp.transform_coordinates(X, np.zeros(6)), X)


def test_quantize_8_3():
    q = osp.quantize_8_3([[1.23456, -0.0004, 12.3454]])
    assert q.dtype == np.float32
    np.testing.assert_allclose(q, [[1.235, -0.0, 12.345]], atol=1e-6)
