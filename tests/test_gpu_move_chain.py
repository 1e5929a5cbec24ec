"""GPU parity of the batched rigid move of the moving chain (gis_move_chain).

The kernel replaces, per candidate pose, transform_coordinates
(pdb_text.py:123-142: (x - mean) @ R + mean + t) and the PDB 8.3 write /
re-read of the moved chain (Z3 pdb_text.py:179-198).  Checked against:

* the oracle's literal text round trip ``float(f"{v:8.3f}")`` on values whose
  product with 1000 is an exact tie (m / 16, m odd: ties to even), their float32
  neighbours and random values (identity pose, zero centre: the rounding alone);
* the host path of the batched step (``RigidBody.moved``: the same rotation
  matrix, NumPy's float64 matmul, the vectorised 8.3 rounding), bit for bit, on
  random poses of a 3,000-atom chain;
* the oracle's transform (its own Rodrigues form of the matrix, whose last bits
  may differ) followed by the text round trip: equal except where the float64
  value sits on a rounding boundary, and then by one unit of the last decimal.
"""
import numpy as np
import pytest
import torch

from oracle import splat as osp

pytestmark = pytest.mark.gpu


def _move(xyz, centre, rt):
    from gisapr_amd import kernels
    dev = torch.device("cuda", 0)
    x = torch.as_tensor(np.ascontiguousarray(xyz, np.float32)).to(dev)
    r = torch.as_tensor(np.ascontiguousarray(rt, np.float64)).to(dev)
    out = kernels.move_chain(x, centre, r)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_rounding_is_the_text_round_trip_with_ties_to_even():
    rng = np.random.default_rng(5)
    m = np.arange(-4001, 4001, 2, dtype=np.float64)
    ties = (m / 16.0).astype(np.float32)  # v * 1000 = 62.5 m: exact ties
    up = np.nextafter(ties, np.float32(np.inf))
    dn = np.nextafter(ties, np.float32(-np.inf))
    rnd = rng.uniform(-999.0, 9999.0, 6000).astype(np.float32)
    small = rng.uniform(-0.01, 0.01, 3000).astype(np.float32)
    vals = np.concatenate([ties, up, dn, rnd, small])
    vals = vals[: len(vals) // 3 * 3].reshape(-1, 3)
    identity = np.concatenate([np.eye(3).reshape(9), np.zeros(3)])[None]
    got = _move(vals, np.zeros(3, np.float32), identity)[0]
    want = osp.quantize_8_3(vals.astype(np.float64))
    assert np.array_equal(got, want)
    # the ties in closed form: 62.5 m to the even integer (Python's round), / 1000
    even = np.array([round(62.5 * v) for v in m], np.float64) / 1000.0
    assert np.array_equal(got.reshape(-1)[: len(ties)], even.astype(np.float32))
    assert (even * 1000.0 < 62.5 * m).any() and (even * 1000.0 > 62.5 * m).any()


def _chain():
    from gisapr_amd import synthetic
    coords, _, chains = synthetic.random_walk_chains([9000, 3000], confine=40.0)
    return coords[chains == chains[-1]] + np.float32(123.456)


def test_matches_host_path_bit_for_bit():
    from gisapr_amd.batch import RigidBody
    xyz = _chain()
    body = RigidBody(np.zeros((0, 3), np.float32), np.zeros(0, np.float32), xyz.astype(np.float32),
                     np.full(len(xyz), 6.0, np.float32))
    rng = np.random.default_rng(11)
    poses = np.concatenate([rng.uniform(-180, 180, (48, 3)), rng.uniform(-60, 60, (48, 3))], 1)
    poses[0] = 0.0
    poses[1, :3] = 0.0
    got = _move(body.moving_coords, body.centre, np.stack([body.pose_rt(p) for p in poses]))
    want = np.stack([body.moved(p) for p in poses])
    assert got.shape == want.shape
    assert np.array_equal(got, want)


def test_matches_oracle_transform_and_text_round_trip():
    xyz = _chain()[:600]
    rng = np.random.default_rng(12)
    poses = np.concatenate([rng.uniform(-180, 180, (16, 3)), rng.uniform(-30, 30, (16, 3))], 1)
    from gisapr_amd.batch import RigidBody
    body = RigidBody(np.zeros((0, 3), np.float32), np.zeros(0, np.float32), xyz, np.ones(len(xyz), np.float32))
    got = _move(xyz, body.centre, np.stack([body.pose_rt(p) for p in poses]))
    want = np.stack([osp.quantize_8_3(osp.transform_coordinates(xyz, p)) for p in poses])
    diff = np.abs(got.astype(np.float64) - want)
    assert (diff > 0).mean() < 1e-4
    assert diff.max() <= 0.0011
