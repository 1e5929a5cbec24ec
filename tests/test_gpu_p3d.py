"""GPU parity of the standalone projector (project3d_and_whiten_cuda.py)
against the float64 oracle (oracle/project.py: volume_dft with and without
padding, central_slice, slice_to_projection, divide_by_spectrum,
normalize_plane, project3d).

Tolerances: relative L2 error 2e-5 on projections, spectra and slices
(float32 transforms of length <= 2N), 5e-5 on whitened / normalised images
(unit-std images after two more transforms and a ring division)."""
import sys

import numpy as np
import pytest

from oracle import project as opj

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def rel_l2(got, want):
    return float(np.linalg.norm((got - want).ravel()) / np.linalg.norm(np.asarray(want).ravel()))


def blob_volume(N, seed):
    """Smooth positive density plus a little noise (a model-like volume)."""
    rng = np.random.default_rng(seed)
    zz, yy, xx = np.indices((N, N, N)).astype(np.float64)
    vol = np.zeros((N, N, N))
    for _ in range(6):
        c = rng.uniform(0.3 * N, 0.7 * N, 3)
        s = rng.uniform(0.04 * N, 0.1 * N)
        vol += rng.uniform(0.5, 1.5) * np.exp(-((zz - c[0]) ** 2 + (yy - c[1]) ** 2 + (xx - c[2]) ** 2) / (2 * s * s))
    return (vol + 0.01 * rng.standard_normal((N, N, N))).astype(np.float32)


ROT = [0.0, 37.5, -120.0, 200.0, 15.0]
TILT = [0.0, 62.0, 90.0, 140.5, 179.0]


@pytest.mark.parametrize("N", [32, 64, 96])
def test_volume_dft_unpadded_matches_oracle(hip_device, N):
    import torch
    from gisapr_amd import kernels
    say("volume_dft unpadded", N)
    vol = np.random.default_rng(N + 1).standard_normal((N, N, N)).astype(np.float32)
    got = kernels.volume_dft_natural(kernels.volume_dft(torch.from_numpy(vol).to(hip_device), pad=False), N,
                                     pad=False).cpu().numpy()
    want = np.fft.ifftshift(opj.volume_dft(vol.astype(np.float64), np.float64, pad=False), axes=(0, 1))
    assert got.shape == want.shape == (N, N, N // 2 + 1)
    assert rel_l2(got, want) < 2e-5


@pytest.mark.parametrize("N", [32, 64])
@pytest.mark.parametrize("pad", [True, False])
@pytest.mark.parametrize("whiten", [False, True])
def test_project_slices_match_oracle(hip_device, N, pad, whiten):
    import torch
    from gisapr_amd import host, kernels
    say("project_slices", N, "pad", pad, "whiten", whiten)
    vol = blob_volume(N, 3)
    dft = kernels.volume_dft(torch.from_numpy(vol).to(hip_device), pad=pad)
    rts = torch.from_numpy(np.ascontiguousarray(host.rotation_matrices_t(ROT, TILT), np.float32)).to(hip_device)
    got = kernels.project_slices(dft, rts, whiten=whiten, normalize=whiten, pad=pad).cpu().numpy()
    want = opj.project3d(vol.astype(np.float64), ROT, TILT, whitening=whiten, pad=pad, dtype=np.float64)
    assert got.shape == want.shape == (len(ROT), N, N)
    for m in range(len(ROT)):
        assert rel_l2(got[m], want[m]) < (5e-5 if whiten else 2e-5), m


@pytest.mark.parametrize("mode", ["divide", "normalize", "both"])
def test_whiten_images_match_oracle(hip_device, mode):
    import torch
    from gisapr_amd import kernels
    say("whiten_images", mode)
    N = 64
    rng = np.random.default_rng(12)
    imgs = np.stack([rng.standard_normal((N, N)).cumsum(0).cumsum(1) * 1e-2 + 0.3 * i for i in range(3)])
    imgs = imgs.astype(np.float32)
    w, z = mode in ("divide", "both"), mode in ("normalize", "both")
    got = kernels.whiten_images(torch.from_numpy(imgs).to(hip_device), whiten=w, normalize=z).cpu().numpy()
    for i, im in enumerate(imgs.astype(np.float64)):
        if mode == "divide":
            want = opj.divide_by_spectrum(im, np.float64)
        elif mode == "normalize":
            want = opj.normalize_plane(im, np.float64)
        else:
            want = opj.whiten(im, np.float64)
        assert rel_l2(got[i], want) < 5e-5, i
    # a pure plane has std 0 < eps: the output is (image - plane) - mean = 0
    yy, xx = np.indices((N, N)).astype(np.float32)
    flat = torch.from_numpy(0.01 * yy - 0.02 * xx + 3.0).to(hip_device)
    assert float(kernels.whiten_images(flat, whiten=False, normalize=True, eps=1e-3).abs().max()) < 1e-4


@pytest.mark.parametrize("L", [64, 96])
def test_extract_central_slices_rfft_matches_oracle(hip_device, L):
    """The reference-layout spectrum (planes and rows fftshifted) sampled at
    rotated central slices equals the oracle's trilinear border sampling;
    zyx matrices give the same slices as their xyz equivalent."""
    import torch
    from gisapr_amd import host, p3d
    say("extract_central_slices_rfft", L)
    vol = blob_volume(L // 2, 5)
    ref = opj.volume_dft(vol.astype(np.float64), np.float64)  # [L, L, L/2+1], fftshifted over z, y
    assert ref.shape == (L, L, L // 2 + 1)
    R = np.ascontiguousarray(host.rotation_matrices_t(ROT, TILT), np.float32)
    dft = torch.from_numpy(ref.astype(np.complex64)).to(hip_device)
    got = p3d.extract_central_slices_rfft(dft, (L, L, L), torch.from_numpy(R).to(hip_device), False).cpu().numpy()
    zyx = p3d.extract_central_slices_rfft(dft, (L, L, L), torch.from_numpy(R[:, ::-1, ::-1].copy()).to(hip_device),
                                          True).cpu().numpy()
    np.testing.assert_array_equal(got, zyx)
    for m, rt in enumerate(R):
        want = opj.central_slice(ref, rt, np.float64)
        assert rel_l2(got[m], want) < 2e-5, m


def test_library_functions_on_single_images(hip_device):
    import torch
    from gisapr_amd import p3d
    say("p3d library functions")
    N = 32
    rng = np.random.default_rng(14)
    im = rng.standard_normal((N, N)).cumsum(1).astype(np.float32)
    t = torch.from_numpy(im).to(hip_device)
    got = p3d.getSpectrum_divideBySpectrum_torch(t, radius_map=p3d.build_radius_map_torch(N, N, hip_device))
    assert rel_l2(got.cpu().numpy(), opj.divide_by_spectrum(im.astype(np.float64), np.float64)) < 5e-5
    yy, xx = p3d.build_xy_grid_torch(N, N, hip_device)
    got = p3d.normalize_torch(t, yy, xx)  # the projector's own (swapped) call
    assert rel_l2(got.cpu().numpy(), opj.normalize_plane(im.astype(np.float64), np.float64)) < 5e-5
    with pytest.raises(ValueError):
        p3d.getSpectrum_divideBySpectrum_torch(t, radius_map=torch.zeros((N, N), dtype=torch.long))


@pytest.mark.parametrize("skip_padding,whiten", [(False, True), (True, False)])
def test_project3d_cli_end_to_end(hip_device, tmp_path, monkeypatch, skip_padding, whiten):
    """MRC model + RELION 3.1 STAR in, <o>.mrcs + <o>.star out (in the working
    directory, as the reference writes them): the stack equals the oracle's
    projections, the STAR keeps the header and every column but the image
    name, which becomes '<row>@<o>.mrcs'."""
    from gisapr_amd import mrc, p3d
    say("p3d CLI", "skip_padding", skip_padding, "whiten", whiten)
    N = 32
    vol = blob_volume(N, 9)
    mrc.write_mrc(tmp_path / "model.mrc", vol)
    header = ["\n", "# version 30001\n", "\n", "data_optics\n", "\n", "loop_\n", "_rlnOpticsGroup #1\n",
              "_rlnImagePixelSize #2\n", "1 1.5\n", "\n", "# version 30001\n", "\n", "data_particles\n", "\n",
              "loop_\n", "_rlnImageName #1\n", "_rlnAngleRot #2\n", "_rlnAngleTilt #3\n", "_rlnAnglePsi #4\n",
              "_rlnOpticsGroup #5\n"]
    rows = [f"{i + 1}@dummy.mrcs {r:.3f} {t:.3f} 0.000 1\n" for i, (r, t) in enumerate(zip(ROT, TILT))]
    (tmp_path / "ang.star").write_text("".join(header + rows))
    monkeypatch.chdir(tmp_path)
    argv = ["--i", "model.mrc", "--ang", "ang.star", "--o", "sub/out.star"]
    if skip_padding:
        argv.append("--skip_padding")
    if whiten:
        argv.append("--normal_background_powerspectrum")
    p3d.main(argv)
    stack = mrc.read(tmp_path / "out.mrcs")
    want = opj.project3d(vol.astype(np.float64), ROT, TILT, whitening=whiten, pad=not skip_padding,
                         dtype=np.float64)
    assert stack.shape == want.shape
    for m in range(len(ROT)):
        assert rel_l2(stack[m], want[m]) < (5e-5 if whiten else 2e-5), m
    out = (tmp_path / "out.star").read_text().splitlines(keepends=True)
    assert out[:len(header)] == header
    body = out[len(header):]
    assert len(body) == len(rows)
    for i, (line, src) in enumerate(zip(body, rows)):
        tok, stok = line.rstrip("\n").split("\t"), src.split()
        assert tok[0] == f"{i + 1}@out.mrcs" and tok[1:] == stok[1:]


@pytest.mark.parametrize("return_array", [True, False])
def test_run_projection_library_form(hip_device, tmp_path, monkeypatch, return_array):
    """project3d_and_whiten_cuda_nowrite.run_projection: the STAR file goes to
    output_filename as given, the stack to output_filename + '.mrcs' unless
    it is returned."""
    from gisapr_amd import mrc, p3d
    say("run_projection", "return_array", return_array)
    N = 32
    vol = blob_volume(N, 10)
    mrc.write_mrc(tmp_path / "model.mrc", vol)
    header = ["data_\n", "loop_\n", "_rlnImageName #1\n", "_rlnAngleRot #2\n", "_rlnAngleTilt #3\n",
              "_rlnAnglePsi #4\n"]
    rows = [f"{i + 1}@x.mrcs {r:.3f} {t:.3f} 10.0\n" for i, (r, t) in enumerate(zip(ROT, TILT))]
    (tmp_path / "ang.star").write_text("".join(header + rows))
    monkeypatch.chdir(tmp_path)
    got = p3d.run_projection(particle_star="ang.star", mrc_volume="model.mrc", whitening=True,
                             do_skip_padding=False, do_return_array=return_array, gpuid=0,
                             output_filename="proj")
    want = opj.project3d(vol.astype(np.float64), ROT, TILT, whitening=True, pad=True, dtype=np.float64)
    if return_array:
        assert not (tmp_path / "proj.mrcs").exists()
        stack = got
        assert stack.dtype == np.float32
    else:
        assert got is None
        stack = mrc.read(tmp_path / "proj.mrcs")
    for m in range(len(ROT)):
        assert rel_l2(stack[m], want[m]) < 5e-5, m
    out = (tmp_path / "proj").read_text().splitlines(keepends=True)
    assert out[:len(header)] == header
    assert [ln.split("\t")[0] for ln in out[len(header):]] == [f"{i + 1}@proj.mrcs" for i in range(len(ROT))]


def test_sample_dft_3d_and_slice_geometry_match_oracle(hip_device):
    """sample_dft_3d against the oracle's trilinear border sampler (coordinates
    inside, on and beyond the edges), and the script's own route to a central
    slice (rotated grid -> mirror x < 0 -> array coordinates -> sample ->
    conjugate) against the oracle's central slice."""
    import torch
    from gisapr_amd import host, p3d
    say("sample_dft_3d")
    rng = np.random.default_rng(21)
    D, H, W = 12, 10, 7
    vol = (rng.standard_normal((D, H, W)) + 1j * rng.standard_normal((D, H, W))).astype(np.complex64)
    coords = np.stack([rng.uniform(-2, D + 1, 500), rng.uniform(-2, H + 1, 500), rng.uniform(-2, W + 1, 500)], -1)
    coords[:3] = [[0, 0, 0], [D - 1, H - 1, W - 1], [5.5, 4.25, 6.0]]
    got = p3d.sample_dft_3d(torch.from_numpy(vol).to(hip_device),
                            torch.from_numpy(coords.astype(np.float32)).to(hip_device)).cpu().numpy()
    want = opj.trilinear_border(vol.astype(np.complex128), coords[:, 0], coords[:, 1], coords[:, 2])
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-5 * np.abs(want).max())
    L = 32
    ref = opj.volume_dft(blob_volume(L // 2, 4).astype(np.float64), np.float64)  # fftshifted rows and planes
    dft = torch.from_numpy(ref.astype(np.complex64)).to(hip_device)
    R = np.ascontiguousarray(host.rotation_matrices_t(ROT, TILT), np.float32)
    grid = p3d.rotated_central_slice_grid((L, L, L), torch.from_numpy(R).to(hip_device), False, rfft=True,
                                          fftshift=True, device=hip_device)
    conj = grid[..., 2] < 0
    grid = torch.where(conj[..., None], -grid, grid)
    s = p3d.sample_dft_3d(dft, p3d.fftfreq_to_dft_coordinates(grid, (L, L, L), rfft=True))
    s = torch.where(conj, torch.conj(s), s).cpu().numpy()
    for m in range(len(ROT)):
        assert rel_l2(s[m], opj.central_slice(ref, R[m], np.float64)) < 2e-5, m
