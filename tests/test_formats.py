"""STAR and MRC formats around the path (host only)."""
import numpy as np
import pytest

from gisapr_amd import mrc, star


def test_mrc_round_trip(tmp_path):
    a = np.random.default_rng(0).standard_normal((3, 5, 7)).astype(np.float32)
    p = tmp_path / "x.mrcs"
    mrc.write_mrc(p, a, voxel_size=1.5)
    h = mrc.read_header(p)
    assert (h["nx"], h["ny"], h["nz"], h["mode"]) == (7, 5, 3, 2)
    np.testing.assert_array_equal(mrc.read(p), a)
    mrc.write_mrc(p, a[0])
    np.testing.assert_array_equal(mrc.mmap(p), a[0])


def test_star_loops_labels_and_selection(tmp_path):
    p = tmp_path / "p.star"
    p.write_text(
        "# relion 3.1\n\ndata_optics\n\nloop_\n_rlnOpticsGroup #1\n_rlnVoltage #2\n1 300\n\n"
        "data_particles\n\nloop_\n_rlnAngleTilt #3\n_rlnImageName #1\n_rlnAngleRot #2\n"
        "_rlnDefocusU #4\n_rlnDefocusV #5\n_rlnDefocusAngle #6\n_rlnAnglePsi #7\n"
        "1@s.mrcs 10.0 20.0 15000 14000 30 5.0\n\n2@s.mrcs 11.0 21.0 16000 15000 31 6.0\nshort row\n")
    t = star.read_particle_table(str(p))
    assert t.has_ctf and len(t.records) == 2
    r = t.records[0]
    assert (r.image_name, r.rot, r.tilt, r.psi, r.defocus_u, r.defocus_angle) == ("1@s.mrcs", 10.0, 20.0, 5.0,
                                                                                15000.0, 30.0)
    lines = [rec.line_index for rec in t.records]
    assert star.select_particle_records(t, lines[1], -1) == [t.records[1]]
    assert star.select_particle_records(t, -1, lines[1]) == [t.records[0]]
    with pytest.raises(ValueError):
        star.find_star_loop(t.lines, ["_rlnMissing"])


def test_direction_star_round_trip(tmp_path):
    p = tmp_path / "a.star"
    star.write_direction_star(p, [-180.0, 12.5], [0.0, 90.0])
    a = star.read_model_angles(str(p))
    assert a.rot == [-180.0, 12.5] and a.tilt == [0.0, 90.0]
    assert a.rotation_matrices_t.shape == (2, 3, 3) and a.rotation_matrices_t.dtype == np.float32


def test_star_loop_selection_rules(tmp_path):
    """v606:96-152 selection: the first loop with the required labels (a loop
    without them is passed over), #n columns or label order, comments and
    blank lines skipped, a loop_ / label line before the first row passed
    over, and rows ending at the next label once one was read."""
    lines = ["data_x\n", "loop_\n", "_rlnAngleRot\n", "_rlnOther\n", "1 2\n",
             "loop_\n", "# c\n", "_rlnAngleTilt #2\n", "\n", "_rlnAngleRot #1\n", "_rlnX\n",
             "loop_\n", "_rlnY\n", "# before rows\n",
             "10 20 x\n", "# mid\n", "11 21 y\n", "_rlnZ\n", "12 22 z\n"]
    loop = star.find_star_loop(lines, ["_rlnAngleRot", "_rlnAngleTilt"])
    assert loop.labels == {"_rlnAngleTilt": 1, "_rlnAngleRot": 0, "_rlnX": 2}
    assert loop.data_line_indices == [14, 16] and loop.data_start == 14
    with pytest.raises(ValueError):
        star.find_star_loop(["loop_\n", "_rlnAngleRot\n", "_rlnAngleTilt\n", "data_y\n"],
                            ["_rlnAngleRot", "_rlnAngleTilt"])


def test_particle_rows_need_every_used_column(tmp_path):
    """Rows too short for the image / angle (or, with defocus labels, the
    defocus) columns are skipped; psi defaults to 0 when its column is absent."""
    p = tmp_path / "q.star"
    p.write_text("data_\nloop_\n_rlnImageName #1\n_rlnAngleRot #2\n_rlnAngleTilt #3\n_rlnAnglePsi #4\n"
                 "1@a 1 2 3\n2@a 4 5\n3@a 6\n")
    t = star.read_particle_table(str(p))
    assert not t.has_ctf
    assert [(r.image_name, r.rot, r.tilt, r.psi) for r in t.records] == [("1@a", 1.0, 2.0, 3.0),
                                                                         ("2@a", 4.0, 5.0, 0.0)]
