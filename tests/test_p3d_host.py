"""Host-side pieces of the standalone projector (no device calls): CLI
flags, the Euler matrix (against the oracle's restatement of
project3d_and_whiten_cuda.py:700-722), file naming and the index maps."""
import numpy as np


def test_parser_flags_match_the_reference():
    from gisapr_amd import p3d
    a = p3d.create_project3d_parser().parse_args(["--i", "m.mrc", "--ang", "a.star", "--o", "o.star"])
    assert (a.i, a.ang, a.o, a.skip_padding, a.normal_background_powerspectrum, a.gpuid) == \
        ("m.mrc", "a.star", "o.star", False, False, None)
    b = p3d.create_project3d_parser().parse_args(["--i", "m", "--ang", "a", "--o", "o", "--skip_padding",
                                                  "--normal_background_powerspectrum", "--gpuid", "0"])
    assert b.skip_padding and b.normal_background_powerspectrum and b.gpuid == "0"


def test_euler_matrix_equals_oracle_and_is_orthonormal():
    from gisapr_amd import p3d
    from oracle import project as opj
    rng = np.random.default_rng(3)
    for a, b, g in rng.uniform(-360, 360, (20, 3)):
        m = np.array(p3d.Euler_angles2matrix(a, b, g))
        np.testing.assert_allclose(m, np.array(opj.euler_angles_to_matrix(a, b, g)), atol=1e-15)
        np.testing.assert_allclose(m @ m.T, np.eye(3), atol=1e-9)


def test_extract_filename_and_index_maps():
    import torch
    from gisapr_amd import p3d
    assert p3d.extract_filename("/a/b/out.star") == "out"
    assert p3d.extract_filename("x.y.star") == "x.y"
    assert p3d.extract_filename("plain") == "plain"
    r = p3d.build_radius_map_torch(8, 8, "cpu").numpy()
    y, x = np.indices((8, 8))
    np.testing.assert_array_equal(r, np.round(np.sqrt((y - 4) ** 2 + (x - 4) ** 2)).astype(np.int64))
    yy, xx = p3d.build_xy_grid_torch(4, 6, "cpu")
    assert torch.equal(yy, torch.arange(4.0)[:, None].expand(4, 6)) and torch.equal(xx, torch.arange(6.0).expand(4, 6))

