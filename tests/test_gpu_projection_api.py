"""GPU parity of the grid_sample projector pieces in the search namespace
(v606:837-1031): the sinc^2 window and the rotated sampling grid against the
float64 oracle (oracle/project.py: sinc2_filter, grid_sample_coordinates),
the whitening against divide_by_spectrum + normalize_plane, a projector
assembled from them as the batched engine assembles it, and the restricted
template build (orientation_indices).

Tolerances: 2e-6 relative on the window (float32 sin / sqrt / division),
2e-6 absolute on normalised grid coordinates (|g| <= ~1.1, float32 products),
5e-5 relative L2 on whitened images, 2e-5 on sampled slices."""
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


@pytest.mark.parametrize("shape", [(64, 64, 64), (9, 10, 13), (8, 12, 20), (3, 1, 2)])
def test_sinc2_filter_matches_oracle(hip_device, shape):
    import torch
    from gisapr_amd.search import apply_sinc2_filter_inplace_3d
    say("sinc2 filter", shape)
    vol = np.random.default_rng(1).standard_normal(shape).astype(np.float32)
    t = torch.from_numpy(vol).to(hip_device)
    assert apply_sinc2_filter_inplace_3d(t, z_chunk=8) is None
    got = t.cpu().numpy()
    want = opj.sinc2_filter(vol.astype(np.float64))
    np.testing.assert_allclose(got, want, rtol=2e-6, atol=1e-30)
    with pytest.raises(ValueError):
        apply_sinc2_filter_inplace_3d(torch.zeros((4, 4), device=hip_device))
    with pytest.raises(ValueError):
        apply_sinc2_filter_inplace_3d(torch.zeros((4, 4, 8), device=hip_device)[:, :, ::2])


@pytest.mark.parametrize("L", [64, 24])
def test_rotated_grid_matches_oracle(hip_device, L):
    import torch
    from gisapr_amd.search import build_projection_base_grid, rotated_grid_to_grid_sample
    say("slice sample grid", L)
    img, dsh = (L, L, L), (L, L, L // 2 + 1)
    base = build_projection_base_grid(img, hip_device)
    rng = np.random.default_rng(L)
    R = np.stack([np.eye(3)] + [opj.euler_angles_to_matrix(*rng.uniform(-180, 180, 3)) for _ in range(4)])
    R = R.astype(np.float32)
    g, conj = rotated_grid_to_grid_sample(base, torch.from_numpy(R).to(hip_device), img, dsh)
    assert g.shape == (5, L, L // 2 + 1, 3) and conj.shape == (5, L, L // 2 + 1) and conj.dtype == torch.bool
    want, wconj = opj.grid_sample_coordinates(base.cpu().numpy(), R, img, dsh)
    np.testing.assert_allclose(g.cpu().numpy(), want, rtol=0, atol=2e-6)
    # the mask agrees wherever the rotated x frequency is not within rounding of 0
    rx = np.einsum("mij,...j->m...i", R.astype(np.float64), base.cpu().numpy()[..., ::-1].astype(np.float64))[..., 0]
    far = np.abs(rx) > 1e-6
    np.testing.assert_array_equal(conj.cpu().numpy()[far], wconj[far])


def test_projection_whitening_matches_oracle(hip_device):
    import torch
    from gisapr_amd.search import ProjectionWhitening
    say("projection whitening")
    N = 64
    rng = np.random.default_rng(5)
    imgs = np.stack([rng.standard_normal((N, N)).cumsum(0).cumsum(1) * 1e-2 + i for i in range(3)]).astype(np.float32)
    pw = ProjectionWhitening(N, hip_device)
    got = pw(torch.from_numpy(imgs).to(hip_device)).cpu().numpy()
    for i in range(3):
        assert rel_l2(got[i], opj.whiten(imgs[i].astype(np.float64), np.float64)) < 5e-5, i
    one = pw(torch.from_numpy(imgs[1:2]).to(hip_device)).cpu().numpy()
    np.testing.assert_array_equal(one[0], got[1])  # batch == 1 is the same path
    with pytest.raises(ValueError):
        pw(torch.from_numpy(imgs[0]).to(hip_device))


def test_assembled_grid_sample_projector_matches_oracle_slices(hip_device):
    """pad -> apply_sinc2_filter_inplace_3d -> fftshift / rfftn / fftshift ->
    rotated_grid_to_grid_sample -> grid_sample -> conj: the central slices of
    the oracle's volume transform."""
    import torch
    import torch.nn.functional as F
    from gisapr_amd.search import apply_sinc2_filter_inplace_3d, build_projection_base_grid, rotated_grid_to_grid_sample
    say("assembled grid_sample projector")
    N = 32
    rng = np.random.default_rng(6)
    vol = rng.standard_normal((N, N, N)).astype(np.float32)
    padded = F.pad(torch.from_numpy(vol).to(hip_device), [N // 2] * 6)
    apply_sinc2_filter_inplace_3d(padded)
    dft = torch.fft.fftshift(torch.fft.rfftn(torch.fft.fftshift(padded, dim=(0, 1, 2)), dim=(0, 1, 2)), dim=(0, 1))
    img, dsh = tuple(padded.shape), tuple(dft.shape)
    R = np.stack([opj.euler_angles_to_matrix(*rng.uniform(-180, 180, 3)) for _ in range(3)]).astype(np.float32)
    grid, conj = rotated_grid_to_grid_sample(build_projection_base_grid(img, hip_device),
                                             torch.from_numpy(R).to(hip_device), img, dsh)
    dft_real = torch.view_as_real(dft).permute(3, 0, 1, 2).unsqueeze(0).contiguous()
    s = F.grid_sample(dft_real, grid.unsqueeze(0), mode="bilinear", padding_mode="border", align_corners=True)
    s = torch.view_as_complex(s.squeeze(0).permute(1, 2, 3, 0).contiguous())
    s = torch.where(conj, torch.conj(s), s).cpu().numpy()
    ref = opj.volume_dft(vol.astype(np.float64), np.float64)
    for m in range(len(R)):
        assert rel_l2(s[m], opj.central_slice(ref, R[m].astype(np.float64), np.float64)) < 2e-5, m


def test_integrated_projector_orientation_indices(hip_device):
    """Restricted projection: rows for the sorted unique indices, identical to
    the same rows of the full bank; out-of-range indices raise."""
    import torch
    from gisapr_amd import search
    from gisapr_amd.star import ModelAngles
    say("IntegratedProjector orientation_indices")
    N, n = 64, 32
    rng = np.random.default_rng(8)
    vol = torch.from_numpy(rng.standard_normal((N, N, N)).astype(np.float32)).to(hip_device)
    rot, tilt = rng.uniform(-180, 180, 7), rng.uniform(0, 180, 7)
    angles = ModelAngles(rot=list(rot), tilt=list(tilt), psi=[0.0] * 7,
                         rotation_matrices_t=opj.rotation_matrices_t(rot, tilt))
    layout = search.PackedFourierLayout(n, 16, hip_device)
    fsc = torch.ones((n, n // 2 + 1), dtype=torch.float32, device=hip_device)
    proj = search.IntegratedProjector(hip_device, N, n, layout, None, fsc)
    full = proj.build_templates(vol, angles)
    assert proj.last_orientation_count == 7
    part = proj.build_templates(vol, angles, [5, 1, 1, 3])
    assert proj.last_orientation_count == 3 and proj.last_slice_count == 3
    assert torch.equal(part, full[[1, 3, 5]])
    assert proj.build_templates(vol, angles, []).shape == (0, n, n // 2 + 1)
    with pytest.raises(ValueError):
        proj.build_templates(vol, angles, [7])
