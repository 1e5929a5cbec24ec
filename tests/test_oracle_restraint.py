"""The geometric-restraint oracle (func_check_boundary_for_testing_v8.py:33-126)
pinned against SciPy: the brute-force distance transform equals
scipy.ndimage.distance_transform_edt, and the float32 pose rotation equals
Rotation.from_rotvec as the reference calls it."""
import numpy as np
import pytest

from oracle import restraint as ors


@pytest.mark.parametrize("shape,p", [((6, 7, 5), 0.05), ((8, 8, 8), 0.01), ((5, 5, 5), 0.5)])
def test_brute_force_edt_equals_scipy(shape, p):
    from scipy.ndimage import distance_transform_edt
    rng = np.random.default_rng(int(p * 100) + shape[0])
    mask = rng.random(shape) < p
    mask[0, 0, 0] = True
    np.testing.assert_allclose(ors.edt_brute_force(mask), distance_transform_edt(~mask), rtol=0, atol=1e-12)


def test_rotation_equals_scipy_float32():
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(2)
    for pose in [[0, 0, 0, 0, 0, 0], [10.0, -20.0, 5.0, 1.0, 2.0, -3.0], *rng.normal(0, 30, (5, 6))]:
        want = Rotation.from_rotvec(np.deg2rad(np.asarray(pose[:3], np.float32))).as_matrix().astype(np.float32)
        np.testing.assert_allclose(ors.rotvec_matrix_f32(pose[:3]), want, atol=2e-7)


def test_restraint_cases():
    """Overlap, separated chain (distance = nearest body voxel), and an empty chain."""
    body = np.array([[0.0, 0.0, 0.0], [1.5, 0.0, 0.0]], np.float32)
    chain = np.array([[9.0, 0.0, 0.0]], np.float32)
    z = np.array([16.0, 16.0])
    ov, d = ors.restraint(chain, [16.0], body, z, [0, 0, 0, -9.0, 0, 0], box=24, apix=1.5)
    assert ov > 0 and d == 0.0
    ov, d = ors.restraint(chain, [16.0], body, z, [0, 0, 0, 0, 0, 0], box=24, apix=1.5)
    assert ov == 0 and d > 0
    ov, d = ors.restraint(chain, [0.0], body, z, [0, 0, 0, 0, 0, 0], box=24, apix=1.5)
    assert ov == 0 and d > 0  # empty chain: distance of (-1, 0, 0) to the body
