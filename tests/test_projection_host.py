"""Host-side tables of the grid_sample projector helpers
(gisapr_amd/projection.py): the base frequency grid and the whitening's ring
map / plane-fit constants, checked against torch's own frequency routines and
the oracle (no GPU needed)."""
import numpy as np
import torch

from oracle import project as opj


def test_base_grid_equals_torch_frequency_routines():
    from gisapr_amd.projection import build_projection_base_grid
    for shape in [(64, 64, 64), (10, 12, 9), (7, 9, 12)]:
        got = build_projection_base_grid(shape, "cpu")
        _, H, W = shape
        fy = torch.fft.fftshift(torch.fft.fftfreq(H, dtype=torch.float32))
        fx = torch.fft.rfftfreq(W, dtype=torch.float32)
        yy, xx = torch.meshgrid(fy, fx, indexing="ij")
        want = torch.stack((torch.zeros_like(yy), yy, xx), dim=-1)
        assert got.dtype == torch.float32 and torch.equal(got, want), shape
        np.testing.assert_allclose(got.numpy(), opj.projection_base_grid(shape), rtol=0, atol=1e-7)


def test_projection_whitening_tables():
    from gisapr_amd.projection import ProjectionWhitening
    n = 64
    pw = ProjectionWhitening(n, "cpu")
    ring = opj.ring_map(n)
    assert torch.equal(pw.radius_map, torch.from_numpy(ring))
    assert pw.max_radius == int(ring.max()) + 1
    np.testing.assert_array_equal(pw.radius_counts.numpy(), np.bincount(ring.ravel()).astype(np.float32))
    yy, xx = np.indices((n, n)).astype(np.float64)
    x, y = yy.ravel(), xx.ravel()  # the reference's swapped names
    want = np.array([[x @ x, x @ y, x.sum()], [x @ y, y @ y, y.sum()], [x.sum(), y.sum(), n * n]])
    np.testing.assert_allclose(pw.plane_matrix.numpy(), want, rtol=1e-7)


def test_search_namespace_resolves_every_batch_engine_name():
    """The names a batched engine looks up in run_search.__globals__."""
    from gisapr_amd import search
    names = ["PackedFourierLayout", "read_model_angles", "read_particle_table", "select_particle_records",
             "build_local_orientation_entries", "generate_soft_mask", "read_fsc_map", "IntegratedProjector",
             "psi_sampling_for_record", "PsiGridBank", "ParticleFourierPreparer", "PackedCTFTorch",
             "PackedCTFNumPy", "ParticleMRCReader", "CorrelationWorkspace", "correlate_templates_and_particle",
             "make_result_records", "chunks", "build_projection_base_grid", "ProjectionWhitening",
             "rotated_grid_to_grid_sample", "apply_sinc2_filter_inplace_3d"]
    g = search.run_search.__globals__
    assert [n for n in names if n not in g] == []
