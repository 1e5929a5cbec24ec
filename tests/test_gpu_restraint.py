"""GPU parity of the geometric restraint (func_check_boundary_for_testing_v8.py:62-126):
the HIP distance transform equals scipy.ndimage.distance_transform_edt exactly,
and GeometryRestraint.evaluate equals the restraint computed with SciPy from
the same device masks (overlap count, minimum distance), for overlapping,
separated and out-of-box chains, at a small box and the reference's default
box of 256."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("N,p", [(48, 0.002), (33, 0.05), (64, 0.0005)])
def test_edt_equals_scipy(hip_device, N, p):
    import torch
    from scipy.ndimage import distance_transform_edt
    from gisapr_amd._lib import check, lib, ptr, stream_of
    rng = np.random.default_rng(N)
    mask = rng.random((N, N, N)) < p
    mask[N // 2, 0, N - 1] = True
    m = torch.from_numpy(mask.astype(np.uint8)).to(hip_device)
    d2 = torch.empty((N, N, N), dtype=torch.float32, device=hip_device)
    say("edt", N, p)
    check(lib().gis_edt_squared(ptr(m), N, ptr(d2), stream_of(hip_device)), "gis_edt_squared")
    got = np.sqrt(d2.cpu().numpy().astype(np.float64))
    np.testing.assert_array_equal(got, distance_transform_edt(~mask))


@pytest.mark.parametrize("box", [64, 256])
def test_geometry_restraint_matches_scipy(hip_device, tmp_path, box):
    import torch
    from scipy.ndimage import distance_transform_edt
    from gisapr_amd import synthetic
    from gisapr_amd.density import density_from_coords
    from gisapr_amd.restraint import GeometryRestraint, rotate_coords_rotvec
    coords, elements, chains = synthetic.random_walk_chains([400, 150], confine=18.0)
    pdb = tmp_path / "m.pdb"
    pdb.write_text(synthetic.pdb_text(coords, elements, chains))
    apix = 1.5
    say("restraint setup", box)
    g = GeometryRestraint(str(pdb), "B", device=hip_device, boxsize=box, apix=apix, threshold=1.0)
    body = g.mainbody_mask.bool().cpu().numpy()
    assert body.any()
    np.testing.assert_array_equal(g.mainbody_distance.cpu().numpy(), distance_transform_edt(~body))
    numbers = g.chain_numbers
    for pose in ([0, 0, 0, 0, 0, 0], [20.0, -10.0, 5.0, 30.0, 0.0, 0.0], [0, 0, 0, 200.0, 0, 0],
                 [5.0, 5.0, 5.0, -25.0, 12.0, 0.0]):
        say("restraint pose", pose)
        ov, dist = g.evaluate(pose)
        xyz = rotate_coords_rotvec(g.chain_coords, *pose)
        chain = (density_from_coords(xyz, numbers, box, apix, 2 * apix, device=hip_device) > 1.0).cpu().numpy()
        want_ov = int(np.count_nonzero(chain & body))
        if want_ov:
            want = (want_ov, 0.0)
        elif chain.any():
            want = (0, float(distance_transform_edt(~body)[chain].min()) * apix)
        else:
            zz, yy, xx = np.nonzero(body)
            want = (0, float(np.sqrt(np.min((zz + 1) ** 2 + yy ** 2 + xx ** 2))) * apix)
        assert (ov, dist) == want, (pose, ov, dist, want)
