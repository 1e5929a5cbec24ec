"""Pins the oracle's projector and whitening (v606:837-1155) with brute force,
the Fourier-slice theorem and invariants."""
import itertools

import numpy as np
import pytest

from oracle import fourier as ofr
from oracle import project as opj


def test_volume_dft_brute_force_tiny():
    N = 4
    rng = np.random.default_rng(0)
    vol = rng.standard_normal((N, N, N))
    got = opj.volume_dft(vol, np.float64)
    L = 2 * N
    pad = np.zeros((L, L, L))
    pad[N // 2:N // 2 + N, N // 2:N // 2 + N, N // 2:N // 2 + N] = vol
    f = (np.arange(L) - N) / L  # fftshift(fftfreq(2N))
    r = np.sqrt(f[:, None, None] ** 2 + f[None, :, None] ** 2 + f[None, None, :] ** 2)
    pad = pad * np.sinc(r) ** 2
    idx = np.arange(L)
    want = np.zeros((L, L, N + 1), complex)
    # direct DFT of the fftshifted array, then fftshift over z and y
    src = np.fft.fftshift(pad)
    zz, yy, xx = np.meshgrid(idx, idx, idx, indexing="ij")
    for kz in range(L):
        for ky in range(L):
            for kx in range(N + 1):
                want[kz, ky, kx] = np.sum(src * np.exp(-2j * np.pi * (kz * zz + ky * yy + kx * xx) / L))
    want = np.fft.fftshift(want, axes=(0, 1))
    np.testing.assert_allclose(got, want, atol=1e-9)


def test_trilinear_border_brute_force():
    rng = np.random.default_rng(1)
    v = rng.standard_normal((5, 6, 4)) + 1j * rng.standard_normal((5, 6, 4))
    pts = np.array([[0.0, 0.0, 0.0], [4.0, 5.0, 3.0], [2.25, 3.5, 1.75], [-1.0, 7.0, 5.0], [3.999, 0.001, 2.5]])
    got = opj.trilinear_border(v, pts[:, 0], pts[:, 1], pts[:, 2])
    for (z, y, x), g in zip(pts, got):
        z, y, x = min(max(z, 0), 4), min(max(y, 0), 5), min(max(x, 0), 3)  # border clamp
        want = 0
        for iz in range(5):
            for iy in range(6):
                for ix in range(4):
                    w = max(0, 1 - abs(z - iz)) * max(0, 1 - abs(y - iy)) * max(0, 1 - abs(x - ix))
                    want += w * v[iz, iy, ix]
        assert abs(g - want) < 1e-12


def test_fourier_slice_theorem_identity_orientation():
    """rot = tilt = 0 samples grid nodes exactly: the projection is the z-sum
    of the sinc^2-weighted padded volume, cropped (no interpolation error)."""
    N = 16
    rng = np.random.default_rng(2)
    vol = rng.standard_normal((N, N, N))
    dft = opj.volume_dft(vol, np.float64)
    rt = opj.rotation_matrices_t([0.0], [0.0])[0]
    proj = opj.slice_to_projection(opj.central_slice(dft, rt, np.float64), np.float64)
    L = 2 * N
    pad = np.pad(vol, N // 2) * opj.sinc2_weights(L, np.float64)
    want = pad.sum(0)[N // 2:N // 2 + N, N // 2:N // 2 + N]
    np.testing.assert_allclose(proj, want, atol=1e-9)


@pytest.mark.parametrize("tilt", [90.0, 270.0])
def test_projection_of_rotated_model_is_rotated_projection(tilt):
    """tilt = +/-90 projects along x, sampling grid nodes only (to 1e-11).
    A band-limited volume (smooth blobs, spectrum ~1e-9 at Nyquist) makes the
    border clamp of the Nyquist row/column irrelevant, so the image equals the
    x-sum of the volume: rows y, columns z (tilt 270) or -z (tilt 90, the
    Friedel-conjugated path, since cos(90 deg) with the literal pi is < 0)."""
    N = 32
    zz, yy, xx = np.indices((N, N, N)).astype(float)
    vol = np.zeros((N, N, N))
    for (cz, cy, cx), a in (((13, 18, 14), 1.0), ((19, 12, 16), 0.7), ((16, 16, 20), 0.5)):
        vol += a * np.exp(-((zz - cz) ** 2 + (yy - cy) ** 2 + (xx - cx) ** 2) / (2 * 2.5 ** 2))
    dft = opj.volume_dft(vol, np.float64)
    rt = opj.rotation_matrices_t([0.0], [tilt])[0]
    proj = opj.slice_to_projection(opj.central_slice(dft, rt, np.float64), np.float64)
    L = 2 * N
    pad = np.pad(vol, N // 2) * opj.sinc2_weights(L, np.float64)
    side = pad.sum(2)  # [z, y] over the padded grid
    view = side.T if tilt == 270.0 else np.roll(side[::-1], 1, axis=0).T  # z -> -z about index N
    want = view[N // 2:N // 2 + N, N // 2:N // 2 + N]
    np.testing.assert_allclose(proj, want, atol=1e-6 * np.abs(want).max())


def test_whitening_invariants():
    rng = np.random.default_rng(5)
    n = 32
    img = rng.standard_normal((n, n)).cumsum(0).cumsum(1)  # strongly coloured
    out = opj.whiten(img, np.float64)
    assert abs(out.mean()) < 1e-10
    assert abs(out.std() - 1.0) < 1e-10
    yy, xx = np.indices((n, n))
    assert abs((out * yy).sum()) < 1e-7 and abs((out * xx).sum()) < 1e-7  # plane removed
    # ring means of |F| are flat before the plane fit: re-derive from the spectrum
    F = np.fft.fftshift(np.fft.fft2(img))
    rm = opj.ring_map(n)
    spec = np.bincount(rm.ravel(), np.abs(F).ravel())
    cnt = np.bincount(rm.ravel())
    div = cnt / spec
    div[0] = 1.0
    G = F * div[rm]
    ring_mean = np.bincount(rm.ravel(), np.abs(G).ravel()) / cnt
    np.testing.assert_allclose(ring_mean[1:], 1.0, rtol=1e-12)


def test_volume_dft_unpadded_brute_force_tiny():
    """--skip_padding (project3d_and_whiten_cuda.py:103-117): no zeros, sinc^2
    over fftshift(fftfreq(N)) of the N grid, then fftshift / rfftn / fftshift."""
    N = 6
    rng = np.random.default_rng(7)
    vol = rng.standard_normal((N, N, N))
    got = opj.volume_dft(vol, np.float64, pad=False)
    f = (np.arange(N) - N // 2) / N
    r = np.sqrt(f[:, None, None] ** 2 + f[None, :, None] ** 2 + f[None, None, :] ** 2)
    src = np.fft.fftshift(vol * np.sinc(r) ** 2)
    idx = np.arange(N)
    zz, yy, xx = np.meshgrid(idx, idx, idx, indexing="ij")
    want = np.zeros((N, N, N // 2 + 1), complex)
    for kz in range(N):
        for ky in range(N):
            for kx in range(N // 2 + 1):
                want[kz, ky, kx] = np.sum(src * np.exp(-2j * np.pi * (kz * zz + ky * yy + kx * xx) / N))
    np.testing.assert_allclose(got, np.fft.fftshift(want, axes=(0, 1)), atol=1e-9)


@pytest.mark.parametrize("pad", [True, False])
def test_project3d_identity_orientation_is_weighted_z_sum(pad):
    """project3d at rot = tilt = 0 samples grid nodes only: the image is the
    z-sum of the sinc^2-weighted (padded) volume, cropped to N."""
    N = 16
    rng = np.random.default_rng(8)
    vol = rng.standard_normal((N, N, N))
    got = opj.project3d(vol, [0.0], [0.0], pad=pad, dtype=np.float64)[0]
    L = 2 * N if pad else N
    o = (L - N) // 2
    full = np.pad(vol, o) * opj.sinc2_weights(L, np.float64)
    np.testing.assert_allclose(got, full.sum(0)[o:o + N, o:o + N], atol=1e-9)


def test_project3d_unpadded_tilt_270_is_weighted_x_sum():
    """Unpadded, tilt 270 projects along x on grid nodes: the image is the
    x-sum of the weighted volume with rows y and columns z (band-limited
    blobs keep the Nyquist border clamp irrelevant)."""
    N = 32
    zz, yy, xx = np.indices((N, N, N)).astype(float)
    vol = np.zeros((N, N, N))
    for (cz, cy, cx), a in (((13, 18, 14), 1.0), ((19, 12, 16), 0.7)):
        vol += a * np.exp(-((zz - cz) ** 2 + (yy - cy) ** 2 + (xx - cx) ** 2) / (2 * 2.5 ** 2))
    got = opj.project3d(vol, [0.0], [270.0], pad=False, dtype=np.float64)[0]
    want = (vol * opj.sinc2_weights(N, np.float64)).sum(2).T
    np.testing.assert_allclose(got, want, atol=1e-6 * np.abs(want).max())


def test_divide_by_spectrum_flattens_every_ring():
    """getSpectrum_divideBySpectrum_torch: the output's own spectrum has ring
    mean |F| = 1 on every ring but DC (whose mean is left as it was)."""
    rng = np.random.default_rng(9)
    n = 32
    img = rng.standard_normal((n, n)).cumsum(0).cumsum(1)
    out = opj.divide_by_spectrum(img, np.float64)
    F = np.fft.fftshift(np.fft.fft2(out))
    rm = opj.ring_map(n)
    cnt = np.bincount(rm.ravel())
    ring_mean = np.bincount(rm.ravel(), np.abs(F).ravel()) / cnt
    np.testing.assert_allclose(ring_mean[1:], 1.0, rtol=1e-10)
    F0 = np.fft.fftshift(np.fft.fft2(img))
    assert abs(ring_mean[0] - np.abs(F0[rm == 0]).mean()) < 1e-9 * abs(ring_mean[0])


def test_normalize_plane_invariants_and_pure_plane():
    """normalize_torch: zero mean, unit population std, no residual plane;
    a pure plane leaves std = 0 < eps and maps to zero."""
    rng = np.random.default_rng(10)
    n = 32
    yy, xx = np.indices((n, n)).astype(float)
    img = 3.0 * yy - 2.0 * xx + 5.0 + rng.standard_normal((n, n))
    out = opj.normalize_plane(img, np.float64)
    assert abs(out.mean()) < 1e-10 and abs(out.std() - 1.0) < 1e-10
    assert abs((out * yy).sum()) < 1e-7 and abs((out * xx).sum()) < 1e-7
    flat = opj.normalize_plane(3.0 * yy - 2.0 * xx + 5.0, np.float64)
    assert np.abs(flat).max() < 1e-9


def test_template_is_unit_norm_and_preshifted():
    N = 16
    rng = np.random.default_rng(6)
    vol = np.abs(rng.standard_normal((N, N, N)))
    mask = opj.generate_soft_mask((N, N), 6, 2)
    fsc = ofr.centered_full_to_packed(opj.expand_radial_values(np.ones(N), N))
    bank = opj.build_templates(vol, [10.0, 100.0], [20.0, 60.0], N, mask, fsc, np.float64)
    for t in bank:
        assert abs(np.linalg.norm(t) - 1.0) < 1e-12
    un = np.fft.ifftshift(bank, axes=(-2, -1))
    # unshifted packed column 0 is the rfft of a real image: Hermitian along rows
    np.testing.assert_allclose(un[:, 1:, 0], np.conj(un[:, :0:-1, 0]), atol=1e-12)


def test_soft_mask_and_radial_expansion_closed_form():
    m = opj.generate_soft_mask((20, 20), 5, 4)
    assert m[10, 10] == 1.0 and m[10, 15] == 1.0 and m[10, 19] == pytest.approx(0.5 * (1 + np.cos(np.pi * 4 / 4)))
    assert m[10, 17] == pytest.approx(0.5 * (1 + np.cos(np.pi * 2 / 4)), abs=1e-7) and m[0, 0] == 0.0
    e = opj.expand_radial_values([1.0, 2.0, 3.0], 8)
    # truncated radius; shells beyond the table are 0
    assert e[4, 4] == 1.0 and e[4, 5] == 2.0 and e[5, 5] == 2.0 and e[4, 6] == 3.0 and e[4, 7] == 0.0
    assert e[6, 6] == 3.0 and e[0, 0] == 0.0  # radius 2.83 -> 2


def test_sinc2_filter_closed_form_on_non_cubic_shapes():
    """Window values at hand-computed frequencies (odd and even axes), and the
    cubic window of the volume transform."""
    import math
    D, H, W = 5, 8, 7
    got = opj.sinc2_filter(np.ones((D, H, W)))
    assert got[2, 4, 3] == 1.0  # zero frequency at n // 2 on every axis

    def s2(r):
        return (math.sin(math.pi * r) / (math.pi * r)) ** 2

    for k in (1, 3):
        assert abs(got[2, 4, 3 + k] - s2(k / 7)) < 1e-15
    r = math.sqrt((2 / 5) ** 2 + (4 / 8) ** 2 + (3 / 7) ** 2)  # corner: -2/5, -1/2, -3/7
    assert abs(got[0, 0, 0] - s2(r)) < 1e-15
    np.testing.assert_allclose(opj.sinc2_filter(np.ones((12, 12, 12))), opj.sinc2_weights(12, np.float64),
                               rtol=0, atol=1e-15)


def test_grid_sample_coordinates_identity_lands_on_the_central_plane():
    """R = I: the align_corners grid indexes voxel (L/2, row j, column kx)
    exactly; no point is mirrored."""
    L = 16
    img, dsh = (L, L, L), (L, L, L // 2 + 1)
    base = opj.projection_base_grid(img)
    assert base.shape == (L, L // 2 + 1, 3) and not base[..., 0].any()
    g, conj = opj.grid_sample_coordinates(base, np.eye(3)[None], img, dsh)
    idx = (g[0] + 1.0) / 2.0 * (np.array(dsh[::-1], np.float64) - 1.0)  # x, y, z voxel indices
    jj, kk = np.meshgrid(np.arange(L), np.arange(L // 2 + 1), indexing="ij")
    np.testing.assert_allclose(idx[..., 0], kk, atol=1e-12)
    np.testing.assert_allclose(idx[..., 1], jj, atol=1e-12)
    np.testing.assert_allclose(idx[..., 2], L // 2, atol=1e-12)
    assert not conj.any()


def test_grid_sample_coordinates_half_turn_mirrors_positive_x():
    """R = diag(-1, -1, 1): every point with fx > 0 is mirrored back onto
    itself; the fx = 0 column keeps -0 (not mirrored) and lands on row L - j."""
    L = 16
    img, dsh = (L, L, L), (L, L, L // 2 + 1)
    base = opj.projection_base_grid(img)
    g0, _ = opj.grid_sample_coordinates(base, np.eye(3)[None], img, dsh)
    g, conj = opj.grid_sample_coordinates(base, np.diag([-1.0, -1.0, 1.0])[None], img, dsh)
    assert conj[0, :, 1:].all() and not conj[0, :, 0].any()
    np.testing.assert_allclose(g[0, :, 1:], g0[0, :, 1:], atol=1e-12)
    yidx = (g[0, :, 0, 1] + 1.0) / 2.0 * (L - 1)
    np.testing.assert_allclose(yidx, (L - np.arange(L)), atol=1e-12)


def test_grid_sample_coordinates_drive_grid_sample_to_the_central_slice():
    """torch's grid_sample (trilinear, border, align_corners) on the oracle
    grid, conjugated by the mask, equals the brute-force-pinned trilinear
    central slice for random rotations."""
    import torch
    import torch.nn.functional as F
    L = 16
    rng = np.random.default_rng(4)
    dft = rng.standard_normal((L, L, L // 2 + 1)) + 1j * rng.standard_normal((L, L, L // 2 + 1))
    R = np.stack([opj.euler_angles_to_matrix(*rng.uniform(-180, 180, 3)) for _ in range(3)]).astype(np.float64)
    img, dsh = (L, L, L), dft.shape
    g, conj = opj.grid_sample_coordinates(opj.projection_base_grid(img), R, img, dsh)
    inp = torch.from_numpy(np.stack((dft.real, dft.imag))[None])
    out = F.grid_sample(inp, torch.from_numpy(g)[None], mode="bilinear", padding_mode="border",
                        align_corners=True)[0].numpy()
    s = out[0] + 1j * out[1]
    s = np.where(conj, np.conj(s), s)
    for m in range(len(R)):
        np.testing.assert_allclose(s[m], opj.central_slice(dft, R[m], np.float64), rtol=0, atol=1e-10)
