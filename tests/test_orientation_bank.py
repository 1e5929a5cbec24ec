"""Calibration banks (gisapr_amd/orientation_bank.py, the zip release's
orientation_bank_v1.py): digest, settings check, row validation and ranking,
on the host.  The digest is pinned against hashlib over the canonical JSON
written out by hand here (sorted keys, no spaces); banks written by the
reference itself are not available (no fixture ships), so cross-reading a
reference-written bank is parity unpinned beyond this format rule."""
import hashlib
import json
import os

import numpy as np
import pytest

from gisapr_amd import orientation_bank as ob


class Rec:
    def __init__(self, line, image):
        self.line_index, self.image_name = line, image


class Angles:
    def __init__(self, rot, tilt):
        self.rot, self.tilt = list(rot), list(tilt)


def test_digest_is_sha256_of_canonical_json_without_digest():
    bank = ob.pack_bank({"schema": ob.SCHEMA, "keep": 2}, {"7": {"image": "1@a.mrcs", "directions": [[3, 10.0, 20.0, 1.5]]}},
                        mrc="/x/m.mrc")
    text = '{"mrc":"/x/m.mrc","rows":{"7":{"directions":[[3,10.0,20.0,1.5]],"image":"1@a.mrcs"}},' \
           '"settings":{"keep":2,"schema":"gisapr-orientations-v1"}}'
    assert bank["sha256"] == hashlib.sha256(text.encode()).hexdigest()
    ob.check_bank(bank, {"schema": ob.SCHEMA, "keep": 2})
    with pytest.raises(ValueError):  # other settings
        ob.check_bank(bank, {"schema": ob.SCHEMA, "keep": 3})
    bank["rows"]["7"]["directions"][0][3] = 1.25  # altered after packing
    with pytest.raises(ValueError):
        ob.check_bank(bank, {"schema": ob.SCHEMA, "keep": 2})


def test_settings_hash_the_input_files(tmp_path):
    (tmp_path / "p.star").write_text("particles\n")
    (tmp_path / "a.star").write_text("angles\n")
    s = ob.settings(tmp_path / "p.star", tmp_path / "a.star", 3, 1, 10, 3.75)
    assert s == {"schema": "gisapr-orientations-v1", "particle_sha256": hashlib.sha256(b"particles\n").hexdigest(),
                 "angle_sha256": hashlib.sha256(b"angles\n").hexdigest(), "keep": 3, "local": True,
                 "local_range": 10.0, "psi_step": 3.75}
    with pytest.raises(Exception):
        ob.positive_int(0)


def test_ranked_row_orders_by_best_psi_with_stable_ties_and_unique_pairs():
    angles = Angles([0.0, 10.0, 20.0, 10.0, 40.0], [5.0, 5.0, 5.0, 5.0, 9.0])
    rec = Rec(4, "2@p.mrcs")
    indices = [0, 1, 2, 3, 4]
    scores = np.array([[1.0, 3.0], [4.0, 2.0], [3.0, 4.0], [5.0, 0.0], [0.5, 0.5]])
    # psi maxima 3, 4, 4, 5, 0.5: direction 3 first; 1 and 2 tie (table order); 3 and 1 share (10, 5)
    row = ob.ranked_row(rec, indices, angles, scores, 3)
    assert row == {"image": "2@p.mrcs", "directions": [[3, 10.0, 5.0, 5.0], [2, 20.0, 5.0, 4.0], [0, 0.0, 5.0, 3.0]]}
    nan = scores.copy()
    nan[2, 1] = np.nan
    assert ob.ranked_row(rec, indices, angles, nan, 3) == {"image": "2@p.mrcs", "directions": [], "invalid_rawcc": True}
    inf = scores.copy()
    inf[0, 0] = np.inf
    with pytest.raises(ValueError):
        ob.ranked_row(rec, indices, angles, inf, 3)
    assert ob.ranked_row(rec, [], angles, np.zeros((0, 2)), 3) == {"image": "2@p.mrcs", "directions": []}


def test_restricted_indices_validates_rows():
    angles = Angles([0.0, 10.0, 20.0], [5.0, 6.0, 7.0])
    rec = Rec(9, "1@p.mrcs")
    good = {"image": "1@p.mrcs", "directions": [[2, 20.0, 7.0, 1.0], [0, 0.0, 5.0, 0.5]]}
    assert ob.restricted_indices(good, rec, [0, 1, 2], angles, 2) == [0, 2]
    bad_rows = [
        None,
        {"image": "2@p.mrcs", "directions": good["directions"]},             # other image
        {"image": "1@p.mrcs", "directions": []},                              # empty while directions exist
        {"image": "1@p.mrcs", "directions": good["directions"] * 2},          # more than keep
        {"image": "1@p.mrcs", "directions": [[1, 10.0, 6.0, 1.0]]},          # not an allowed direction
        {"image": "1@p.mrcs", "directions": [[2, 20.0, 8.0, 1.0]]},          # rot/tilt not the table's
        {"image": "1@p.mrcs", "directions": [[2, 20.0, 7.0, float("nan")]]},  # non-finite score
        {"image": "1@p.mrcs", "directions": [[2.0, 20.0, 7.0, 1.0]]},        # index not an int
    ]
    for row in bad_rows:
        with pytest.raises(ValueError):
            ob.restricted_indices(row, rec, [0, 2], angles, 2)
    assert ob.restricted_indices({"image": "1@p.mrcs", "directions": []}, rec, [], angles, 2) == []


def test_atomic_json_replaces_in_place(tmp_path):
    path = tmp_path / "sub" / "bank.json"
    ob.atomic_json(path, {"a": 1})
    ob.atomic_json(path, {"b": [1, 2]})
    assert json.loads(path.read_text()) == {"b": [1, 2]}
    assert os.listdir(path.parent) == ["bank.json"]
    with pytest.raises(ValueError):
        ob.atomic_json(path, {"x": float("nan")})
    assert json.loads(path.read_text()) == {"b": [1, 2]} and os.listdir(path.parent) == ["bank.json"]
