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



def test_projector_geometry_helpers():
    """The projector script's frequency-grid helpers on the CPU, against numpy
    and closed forms: an unrotated central slice maps to integer array
    coordinates of the fftshifted rfft (plane L/2, row j, column k)."""
    import numpy as np
    import torch
    from gisapr_amd import p3d
    from oracle import project as opj
    assert p3d.rfft_shape((6, 8, 10)) == (6, 8, 6) and p3d.rfft_shape((5, 7)) == (5, 4)
    assert abs(p3d.DEG2RAD(180.0) - 3.14159265359) < 1e-15
    assert p3d.dft_center((8, 8, 8), rfft=True, fftshifted=True).tolist() == [4, 4, 0]
    assert p3d.dft_center((7, 9), rfft=False, fftshifted=True).tolist() == [3, 4]
    assert p3d.dft_center((7, 9), rfft=False, fftshifted=False).tolist() == [0, 0]
    g = p3d.fftfreq_grid((6, 8, 10), rfft=True, spacing=(1.0, 2.0, 0.5)).numpy()
    assert g.shape == (6, 8, 6, 3)
    np.testing.assert_allclose(g[:, 0, 0, 0], np.fft.fftfreq(6, 1.0), atol=1e-7)
    np.testing.assert_allclose(g[0, :, 0, 1], np.fft.fftfreq(8, 2.0), atol=1e-7)
    np.testing.assert_allclose(g[0, 0, :, 2], np.fft.rfftfreq(10, 0.5), atol=1e-7)
    gs = p3d.fftfreq_grid((6, 8), rfft=False, fftshift=True).numpy()
    np.testing.assert_allclose(gs[:, 3, 0], np.fft.fftshift(np.fft.fftfreq(6)), atol=1e-7)
    np.testing.assert_allclose(gs[2, :, 1], np.fft.fftshift(np.fft.fftfreq(8)), atol=1e-7)
    n = p3d.fftfreq_grid((6, 8), rfft=False, norm=True).numpy()
    np.testing.assert_allclose(n, np.hypot(*np.meshgrid(np.fft.fftfreq(6), np.fft.fftfreq(8), indexing="ij")),
                               atol=1e-7)
    L = 16
    base = p3d.central_slice_grid((L, L, L), rfft=True, fftshift=True)
    np.testing.assert_allclose(base.numpy(), opj.projection_base_grid((L, L, L)), atol=1e-7)
    eye = torch.eye(3)[None]
    rot = p3d.rotated_central_slice_grid((L, L, L), eye, True, rfft=True, fftshift=True)
    coords = p3d.fftfreq_to_dft_coordinates(rot[0], (L, L, L), rfft=True).numpy()
    jj, kk = np.meshgrid(np.arange(L), np.arange(L // 2 + 1), indexing="ij")
    np.testing.assert_allclose(coords[..., 0], L // 2, atol=1e-5)
    np.testing.assert_allclose(coords[..., 1], jj, atol=1e-5)
    np.testing.assert_allclose(coords[..., 2], kk, atol=1e-5)
    # xyz and zyx forms of one rotation give the same grid; grid-sample normalisation as the oracle's
    R = torch.tensor(opj.euler_angles_to_matrix(30.0, 50.0, 10.0), dtype=torch.float32)
    a = p3d.rotated_central_slice_grid((L, L, L), R[None], False, rfft=True)
    b = p3d.rotated_central_slice_grid((L, L, L), R.flip(-1).flip(-2)[None], True, rfft=True)
    np.testing.assert_allclose(a.numpy(), b.numpy(), atol=1e-6)
    c = torch.tensor([[[0.0, 0.0, 0.0], [15.0, 7.5, 8.0]]])
    np.testing.assert_allclose(p3d.array_to_grid_sample(c, (16, 16, 9)).numpy(),
                               [[[-1.0, -1.0, -1.0], [1.0, 0.0, 1.0]]], atol=1e-6)


def test_projector_star_helpers():
    from gisapr_amd import p3d
    v31 = ["\n", "# version 30001\n", "\n", "data_optics\n", "loop_\n", "_rlnOpticsGroup #1\n", "1\n", "\n",
           "data_particles\n", "loop_\n", "_rlnImageName #1\n", "_rlnAngleRot #2\n", "1@a.mrcs 10\n"]
    assert p3d.judge_relion30_or_relion31(v31) == 0
    v30 = ["\n", "data_\n", "\n", "loop_\n", "_rlnImageName #1\n", "_rlnAngleRot #2\n", "1@a.mrcs 10\n"]
    assert p3d.judge_relion30_or_relion31(v30) == 1
    assert p3d.judge_relion30_or_relion31(["# RELION optimiser 3.0.8\n", "# version 30001\n", "\n"]) == 1
    assert p3d.judge_mline0(v30) == 6
    assert p3d.judge_mline0(v31) == 6 and p3d.judge_mline1(v31, 7) == 12
    assert p3d.judge_mline0(["data_\n", "loop_\n"]) is None
