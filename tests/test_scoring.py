"""ReSuLt score files (gisapr_amd/scoring.py; new_method_to_fit_the_2nd_Gaussian_PEAK_v4.py,
Z3 score_index :38-58) from search records in the v606 text format
(v606:1494-1516): closed forms of the simple sums, the fit of a planted
two-peak sample, the failed-fit fallback and the optimizer's reading of the
file (func.py:113-127)."""
import math

import numpy as np
import pytest

from gisapr_amd import scoring


def write_records(path, rawcc, z):
    with open(path, "w") as fh:
        for i, (r, s) in enumerate(zip(rawcc, z)):
            fh.write(str(["cc: ", 10 + i, float(s), 1 if s > 3 else 0, 1, 12.0, 34.0, 2.5, 1, -2, 0, 0]) + "\n")
            fh.write(str(["psi: ", 10 + i, 45.0, 47.5]) + "\n")
            fh.write(str(["rawCC: ", 10 + i, float(r)]) + "\n")


def test_simple_sums_and_result_file(tmp_path):
    rng = np.random.default_rng(0)
    raw = np.concatenate([rng.normal(0.15, 0.02, 400), rng.normal(0.35, 0.03, 200)]) * 1024.0
    z = rng.normal(0.0, 1.0, 600)
    write_records(tmp_path / "a.txt", raw, z)
    (tmp_path / "index.txt").write_text("a.txt\n")
    scoring.score_index(str(tmp_path / "index.txt"), 1, 1, str(tmp_path / "ReSuLt_a.txt"), verbose=False,
                        base_dir=str(tmp_path))
    lines = (tmp_path / "ReSuLt_a.txt").read_text().splitlines()
    assert lines[0] == "a.txt"
    total, zv, cc = (float(v) for v in lines[1].split())
    assert total == pytest.approx(float(np.sum(np.exp(raw / 1024.0))), rel=1e-12)   # v4:133-134
    assert zv == 0.0 and raw.min() / 1024.0 < cc < raw.max() / 1024.0  # a fitted mean inside the data
    assert scoring.read_score(tmp_path / "ReSuLt_a.txt") == -total
    # z-score mode with the simple sum: sum of the z values
    _, _, _, tz = scoring.run_this(str(tmp_path / "a.txt"), 0, 1, verbose=False)
    assert tz == pytest.approx(float(np.sum(z)), rel=1e-12)


def test_integrated_estimate_above_peak(tmp_path):
    """Without the simple sum: (count above the peak) x peak + (sum above it)."""
    rng = np.random.default_rng(1)
    raw = np.concatenate([rng.normal(0.1, 0.015, 500), rng.normal(0.3, 0.03, 300)])
    write_records(tmp_path / "b.txt", raw * 1024.0, np.zeros(800))
    _, peak, _, total = scoring.run_this(str(tmp_path / "b.txt"), 1, 0, verbose=False)
    data = raw * 1024.0 / 1024.0
    above = data > peak
    assert total == pytest.approx(above.sum() * peak + data[above].sum(), rel=1e-12)
    assert data.min() < peak < data.max()


def test_empty_and_failed_fit(tmp_path, monkeypatch):
    (tmp_path / "empty.txt").write_text("")
    assert scoring.run_this(str(tmp_path / "empty.txt"), 1, 1, verbose=False) == (0.0, 0.0, 0.0, 0.0)
    write_records(tmp_path / "two.txt", [512.0, 256.0], [1.0, 2.0])
    import scipy.optimize

    def fail(*a, **k):
        raise RuntimeError("Optimal parameters not found")
    monkeypatch.setattr(scipy.optimize, "curve_fit", fail)
    # a failed fit scores the plain sum of the values (rawCC / 1024 or z), peak 0
    assert scoring.run_this(str(tmp_path / "two.txt"), 1, 1, verbose=False) == (0.0, 0.0, 0.0, 0.75)
    assert scoring.run_this(str(tmp_path / "two.txt"), 0, 0, verbose=False) == (0.0, 0.0, 0.0, 3.0)
    (tmp_path / "bad.txt").write_text("x\nnan 0 0\n")
    assert scoring.read_score(tmp_path / "bad.txt") == 0.0
    (tmp_path / "inf.txt").write_text("x\ninf 0 0\n")
    with pytest.raises(RuntimeError):
        scoring.read_score(tmp_path / "inf.txt")
