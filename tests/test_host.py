"""Product host-side setup (gisapr_amd.host) against the oracle restatement."""
import numpy as np
import pytest

from gisapr_amd import host, synthetic
from oracle import fourier as ofr
from oracle import particles as opt
from oracle import project as opj
from oracle import search as osr


def test_rotation_matrices_match_oracle():
    rot, tilt = synthetic.healpix_nest_directions(2)
    np.testing.assert_array_equal(host.rotation_matrices_t(rot, tilt), opj.rotation_matrices_t(rot, tilt))
    e = host.euler_batch(rot[:20], tilt[:20])
    for i in range(20):
        np.testing.assert_allclose(e[i], opj.euler_angles_to_matrix(rot[i], tilt[i], 0.0), atol=1e-15)


def test_local_orientation_map_matches_oracle():
    drot, dtilt = synthetic.healpix_nest_directions(2)
    rng = np.random.default_rng(3)
    pr = rng.uniform(-180, 180, 60)
    pt = np.degrees(np.arccos(rng.uniform(-1, 1, 60)))
    got = host.local_orientation_map(pr, pt, drot, dtilt, 15.0)
    for i in range(60):
        assert got[i] == osr.local_orientations(pr[i], pt[i], drot, dtilt, 15.0)
    assert sum(len(g) for g in got) > 0


@pytest.mark.parametrize("psi,step,local,rng_", [(12.0, 7.5, True, 15.0), (359.9, 3.75, True, 10.0),
                                                 (0.2, 15.0, False, 20.0), (181.875, 3.75, True, 10.0)])
def test_psi_sampling_matches_oracle(psi, step, local, rng_):
    assert host.psi_sampling_for_record(psi, step, local, rng_) == \
           opt.psi_sampling_for_record(psi, step, local, rng_)


def test_psi_grid_params_match_oracle_sampler():
    N = 32
    for a in (0.5, 7.5, -22.5, 183.75):
        c, ms, s, c2, oy, ox = host.psi_grid_params(N, [a])[0]
        yy, xx = np.indices((N, N)).astype(np.float64)
        sy, sx = opt.rotation_source(N, a)
        np.testing.assert_allclose(float(c) * yy + float(ms) * xx + float(oy), sy, atol=1e-5)
        np.testing.assert_allclose(float(s) * yy + float(c2) * xx + float(ox), sx, atol=1e-5)


def test_ctf_tables_reproduce_oracle_ctf():
    n, px, kk = 64, 1.5, 3.0
    tab = host.CTFTables(n, px, 2.7, 300.0, kk)
    octf = opt.PackedCTF(n, px, 2.7, 300.0, kk)
    x, y, u4, nc = tab.table
    np.testing.assert_array_equal(x, octf.x)
    np.testing.assert_array_equal(y, octf.y)
    np.testing.assert_array_equal(u4, octf.u4)
    np.testing.assert_array_equal(nc, octf.ncurve)
    assert np.float32(tab.k1) == octf.k1 and np.float32(tab.k2) == octf.k2 and np.float32(tab.k3) == octf.k3


def test_mask_and_fsc_match_oracle():
    np.testing.assert_array_equal(host.generate_soft_mask((40, 40), 15, 4), opj.generate_soft_mask((40, 40), 15, 4))
    vals = np.linspace(1, 0, 17).astype(np.float32)
    want = ofr.centered_full_to_packed(opj.expand_radial_values(vals, 32))
    np.testing.assert_array_equal(host.packed_fsc(vals, 32), want)


def test_healpix_counts_and_poles():
    for order in (0, 1, 2, 3):
        rot, tilt = synthetic.healpix_nest_directions(order)
        assert len(rot) == 12 * 4 ** order
        assert np.all((rot >= -180) & (rot < 180)) and np.all((tilt >= 0) & (tilt <= 180))
    rot, tilt = synthetic.healpix_nest_directions(0)
    # base resolution: rings at cos(theta) = 2/3, 0, -2/3
    np.testing.assert_allclose(np.sort(np.unique(np.round(np.cos(np.radians(tilt)), 6))), [-2 / 3, 0, 2 / 3],
                               atol=1e-6)
    # directions are distinct unit vectors, roughly uniform: mean vector ~ 0
    rot, tilt = synthetic.healpix_nest_directions(3)
    v = np.stack([np.sin(np.radians(tilt)) * np.cos(np.radians(rot)),
                  np.sin(np.radians(tilt)) * np.sin(np.radians(rot)), np.cos(np.radians(tilt))], 1)
    assert np.linalg.norm(v.mean(0)) < 1e-6
    assert len(np.unique(np.round(v, 9), axis=0)) == len(v)


def test_volume_dft_brick_unpack_is_the_documented_index():
    """kernels.volume_dft_natural inverts the bricked index of gis_volume_dft."""
    import torch
    from gisapr_amd import kernels
    N = 32
    xb = N // 4 + 1
    nat = torch.arange(2 * N * 2 * N * (N + 1), dtype=torch.float32).reshape(2 * N, 2 * N, N + 1)
    bricked = torch.zeros(16 * N * N * xb, dtype=torch.complex64)
    kz, ky, kx = np.meshgrid(np.arange(2 * N), np.arange(2 * N), np.arange(N + 1), indexing="ij")
    idx = ((((kz // 2) * N + ky // 2) * xb + kx // 4) * 2 + kz % 2) * 2 * 4 + (ky % 2) * 4 + kx % 4
    bricked[torch.from_numpy(idx.ravel())] = nat.reshape(-1).to(torch.complex64)
    np.testing.assert_array_equal(kernels.volume_dft_natural(bricked, N).real.numpy(), nat.numpy())


def test_quantize_8_3_matches_text_round_trip():
    """The vectorised 8.3 quantiser equals float(f"{x:8.3f}") on random values,
    exact decimal ties (x.xxx5) and negative zero."""
    from gisapr_amd.pdb_text import quantize_8_3
    rng = np.random.default_rng(7)
    vals = np.concatenate([rng.uniform(-999, 9999, 30000), rng.uniform(-1, 1, 3000),
                           np.arange(-5000, 5000) / 2000.0 + 0.0005, [-0.0001, -0.0004, 0.0005, -0.0005, 2.0005]])
    vals = vals[: (len(vals) // 3) * 3].reshape(-1, 3)
    want = np.asarray([[float(f"{v:8.3f}") for v in r] for r in vals], np.float32)
    got = quantize_8_3(vals)
    np.testing.assert_array_equal(got, want)
    assert np.array_equal(np.signbit(got), np.signbit(want))
    with pytest.raises(ValueError):
        quantize_8_3([[10000.0, 0.0, 0.0]])


def test_pose_operand_of_the_device_move():
    """RigidBody.pose_rt: the rotation matrix ``moved`` applies and the
    translation; poses that could leave the 8.3 field go through the host
    check, which raises for a coordinate that really leaves it."""
    from gisapr_amd.batch import RigidBody
    from gisapr_amd.pdb_text import rotvec_matrix
    xyz = np.array([[9990.0, 0.0, 0.0], [9980.0, 1.0, 2.0], [9985.0, -3.0, 1.0]], np.float32)
    body = RigidBody(np.zeros((0, 3), np.float32), np.zeros(0, np.float32), xyz, np.ones(3, np.float32))
    pose = np.array([10.0, -20.0, 30.0, 1.0, 2.0, 3.0])
    rt = body.pose_rt(pose)
    assert rt.shape == (12,) and rt.dtype == np.float64
    np.testing.assert_array_equal(rt[:9], rotvec_matrix(pose[:3]).reshape(9))
    np.testing.assert_array_equal(rt[9:], pose[3:])
    body.pose_rt([0, 0, 0, 9.9, 0, 0])  # the bound trips (9985 + 9.9 + 5.4), the host check passes (x <= 9999.9)
    with pytest.raises(ValueError):
        body.pose_rt([0, 0, 0, 20.0, 0, 0])
    with pytest.raises(ValueError):
        body.pose_rt([0, 0, np.nan, 0, 0, 0])


def test_search_session_cache_forms():
    """run_search's metadata cache accepts the worker's SearchCache shape
    (``metadata.get_or_build``), a dict, or None."""
    from gisapr_amd.search import _session_get
    calls = []

    def factory():
        calls.append(1)
        return len(calls)

    class Store:
        def __init__(self):
            self.d = {}

        def get_or_build(self, key, f):
            if key not in self.d:
                self.d[key] = f()
            return self.d[key]

    class SearchCacheLike:
        def __init__(self):
            self.metadata = Store()

    sc = SearchCacheLike()
    assert _session_get(sc, ("a", 1), factory) == 1 and _session_get(sc, ("a", 1), factory) == 1
    d = {}
    assert _session_get(d, "k", factory) == 2 and _session_get(d, "k", factory) == 2 and d == {"k": 2}
    assert _session_get(None, "k", factory) == 3 and _session_get(None, "k", factory) == 4


def test_v606_small_helpers_and_local_cache(tmp_path):
    """clip_value, dot_product, the label-column parser, and the local
    orientation cache: built, written, reused; a cache holding anything but
    plain lists and ints is refused (no object is constructed) and rebuilt."""
    import pickle
    import numpy as np
    from gisapr_amd import search
    from gisapr_amd.star import ModelAngles, ParticleRecord
    assert search.clip_value(5, 0, 3) == 3.0 and search.clip_value(-1, 0, 3) == 0.0 and search.clip_value(2, 0, 3) == 2.0
    assert search.dot_product([1, 2], [3, 4]) == 11.0 and search.dot_product([1], [1, 2]) == -9999999.0
    assert search._parse_label_index(["_rlnAngleRot", "#3"], 9) == 2 and search._parse_label_index(["_x"], 9) == 9
    recs = [ParticleRecord(10, "1@a", 0.0, 30.0, 0.0), ParticleRecord(11, "2@a", 90.0, 60.0, 0.0)]
    angles = ModelAngles([0.0, 5.0, 90.0, 180.0], [30.0, 32.0, 60.0, 90.0], [0.0] * 4, np.zeros((4, 3, 3)))
    star_path = str(tmp_path / "p.star")
    m1 = search.load_or_build_local_orientation_map(star_path, recs, angles, 10.0)
    want = {e[0]: e[1:] for e in search.build_local_orientation_entries(recs, angles.rot, angles.tilt, 10.0)}
    assert m1 == want
    cache = tmp_path / "p.star_LocalSearch_Angle_10_data.pkl"
    assert pickle.loads(cache.read_bytes()) == [[10] + want[10], [11] + want[11]]
    assert search.load_or_build_local_orientation_map(star_path, recs, angles, 10.0) == m1  # reused

    class Evil:
        def __reduce__(self):
            return (os.system, ("true",))

    import os
    cache.write_bytes(pickle.dumps([[10, Evil()]]))
    assert search.load_or_build_local_orientation_map(star_path, recs, angles, 10.0) == m1  # refused, rebuilt
    assert pickle.loads(cache.read_bytes()) == [[10] + want[10], [11] + want[11]]


def test_pdb2mrc_cli_helpers():
    import torch
    from gisapr_amd import density
    a = density.create_parser().parse_args(["--i", "m.pdb", "--o", "m.mrc", "--box", "64", "--center",
                                            "--dontWriteMRC"])
    assert (a.box, a.center, a.dontWriteMRC, a.gpuid, a.device) == (64, True, True, 0, None)


@pytest.mark.parametrize("n", [4, 32, 40, 64, 96, 160, 192, 256])
def test_fragment_orders_are_permutations(n):
    """Both operand orders place every packed coefficient at a distinct slot
    of the plane; the fp16 order keeps a lane's four rows of one column
    contiguous and the same tiles and lone column as the float32 order."""
    from gisapr_amd import kernels
    m = n // 2 + 1
    nks = ((n + 15) // 16 + 1) & ~1
    lone = m % 16 == 1
    nct = m // 16 if lone else (m + 15) // 16
    F = nks * nct * 256 + (nks * 16 if lone else 0)
    for pos in (kernels.fragment_positions(n), kernels.half_fragment_positions(n)):
        assert pos.shape == (n, m)
        assert len(np.unique(pos)) == n * m and pos.min() >= 0 and pos.max() < F
    p32, p16 = kernels.fragment_positions(n), kernels.half_fragment_positions(n)
    np.testing.assert_array_equal(p32 // 256, p16 // 256)  # same tile (and lone column) of every value
    r, c = np.meshgrid(np.arange(n), np.arange(m), indexing="ij")
    inside = p16 < nks * nct * 256
    lane = (((r >> 2) & 3) * 16 + (c & 15))
    np.testing.assert_array_equal((p16 % 256)[inside], (lane * 4 + (r & 3))[inside])
