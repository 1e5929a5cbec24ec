"""Pins the oracle's records and scores (v606:331-378, 1452-1516; v4:133-134)."""
import math

import numpy as np
import pytest

from oracle import search as osr


def test_angular_distance_closed_forms():
    # pure rot change at tilt 0: two row vectors turn by d, the third stays -> mean 2d/3
    assert osr.calculate_angular_distance(10.0, 0.0, 0.0, 40.0, 0.0, 0.0) == pytest.approx(20.0, abs=1e-6)
    assert osr.calculate_angular_distance(33.0, 71.0, 0.0, 33.0, 71.0, 0.0) == pytest.approx(0.0, abs=1e-5)
    # pure tilt change at rot 0: rows 0 and 2 turn by d, row 1 stays
    assert osr.calculate_angular_distance(0.0, 10.0, 0.0, 0.0, 25.0, 0.0) == pytest.approx(10.0, abs=1e-6)


def test_local_orientations_rule():
    rot = np.array([0.0, 5.0, 14.0, 16.0, 180.0, -175.0])
    tilt = np.array([90.0, 90.0, 90.0, 90.0, 90.0, 90.0])
    # prefilter |drot|, |dtilt| <= r, then mean axis angle <= r
    got = osr.local_orientations(178.0, 90.0, rot, tilt, 10.0)
    assert got == [4, 5]  # wrap-around at +/-180
    got = osr.local_orientations(0.0, 90.0, rot, tilt, 15.0)
    assert got == [0, 1, 2]


def test_make_record_first_max_zscore_shift():
    s = np.array([[1.0, 3.0], [3.0, 0.5]], np.float32)  # tie: first max (0, 1) wins
    ys = np.array([[0, 5], [1, 2]])
    xs = np.array([[0, 7], [1, 2]])
    rec = osr.make_record(7, (10.0, 20.0, 30.0), [100.0, 107.5], [10.0, 12.0], [20.0, 22.0], [0, 1], s, ys, xs,
                          crop=8, trans_scale=1.0)
    assert (rec["best_local"], rec["best_psi"]) == (0, 1)
    assert rec["rawcc"] == 3.0 and rec["psi_fake"] == 107.5
    mean = s.mean(dtype=np.float64)
    assert rec["z"] == pytest.approx((3.0 - mean) / s.std(dtype=np.float64))
    assert rec["xshift"] == 8 // 2 - 7 and rec["yshift"] == 8 // 2 - 5
    # shift uses Python round (half-to-even) after dividing by trans_scale
    rec = osr.make_record(7, (10.0, 20.0, 30.0), [0.0, 0.0], [10.0], [20.0], [0], s[:1], ys[:1], xs[:1], crop=8,
                          trans_scale=2.0)
    assert rec["xshift"] == round((4 - 7) / 2.0) == -2


def test_score_sum_closed_form():
    raw = [1024.0, 0.0, -2048.0]
    assert osr.score_sum(raw) == pytest.approx(math.e + 1.0 + math.exp(-2.0))
