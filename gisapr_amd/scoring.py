"""Conformation scores from the per-particle search records, and the
``ReSuLt_*`` score file the optimizers read
(new_method_to_fit_the_2nd_Gaussian_PEAK_v4.py; Z3 score_index :38-58).

A search writes three text records per particle (v606:1494-1516),
``['cc: ', line, z, l3, accu, rot, tilt, angdist, xshift, yshift, 0, 0]``,
``['psi: ', ...]`` and ``['rawCC: ', line, peak]``.  A score file lists, for
each result file named in an index file, the file name and a line
"<total> <zscore value> <cc value>"; the optimizer takes <total> from the
second line (func.py:113-127).

The score of one result file, as the reference computes it (the same steps,
so the same files give the same numbers):
  * CC mode (do_run_CC > 0), over rawCC / 1024 of every particle: a
    1000-bin density histogram on [0, 1], smoothed by a Gaussian of sigma 2
    bins, fitted by two Gaussians (scipy curve_fit from the initial guess
    [h, x_peak, 0.02, h / 2, x_peak + 0.05, 0.1]); the refined peak is the
    larger of the two fitted means; total = (number of values above it) x
    peak + (sum of the values above it); with do_simple_sum, total =
    sum exp(rawCC / 1024).  A failed fit gives total = sum of the values.
  * z-score mode (do_run_CC <= 0), over the z values below 9999: the same
    with a 1000-bin histogram over the data's range, initial guess
    [h, x_peak, 0.3, h / 4, x_peak + 2, 4], and, with do_simple_sum, total =
    sum of the z values.
The returned tuple is (0, refined peak, 0, total); a file without records
scores zeros.  The in-memory equivalent of the simple sum on the device is
kernels.score_sum (float64).
"""
from __future__ import annotations

from pathlib import Path
from typing import List, Tuple

import numpy as np


def _gauss(x, amp, mean, sd):
    return amp * np.exp(-((x - mean) ** 2) / (2 * sd ** 2))


def _two_gauss(x, a1, m1, s1, a2, m2, s2):
    return _gauss(x, a1, m1, s1) + _gauss(x, a2, m2, s2)


def read_result_records(path) -> Tuple[List[float], List[float]]:
    """(z values below 9999, rawCC / 1024 values) of one search result file."""
    zscores, raw = [], []
    with open(path, "r") as fh:
        for line in fh:
            fields = line.split(",")
            tag = fields[0]
            if tag == "['cc: '":
                z = float(fields[2])
                if z < 9999:
                    zscores.append(z)
            elif tag == "['rawCC: '":
                raw.append(float(fields[-1].split("]")[0]) / 1024.0)
    return zscores, raw


def _peak_fit(values: np.ndarray, hist_range, guess_width, guess_second):
    """The refined peak (the larger fitted mean), or None when the fit fails."""
    from scipy.ndimage import gaussian_filter
    from scipy.optimize import curve_fit
    hist, edges = np.histogram(values, bins=1000, range=hist_range, density=True)
    smooth = gaussian_filter(hist, sigma=2)
    centres = (edges[:-1] + edges[1:]) / 2
    h = max(smooth)
    x0 = centres[int(np.argmax(smooth))]
    guess = [h, x0, guess_width, h * guess_second[0], x0 + guess_second[1], guess_second[2]]
    try:
        params, _ = curve_fit(_two_gauss, centres, smooth, p0=guess)
    except Exception:  # the reference catches every fit failure
        return None
    return params[1] if params[1] > params[4] else params[4]  # the larger fitted mean


def run_this(results_file, do_run_CC, do_simple_sum, verbose=True):
    """Score of one result file: (0.0, refined peak, 0.0, total)."""
    zscores, raw = read_result_records(results_file)
    if len(raw) < 1:
        return 0.0, 0.0, 0.0, 0.0
    if do_run_CC > 0:
        data = np.asarray(raw)
        peak = _peak_fit(data, (0, 1.0), 0.02, (0.5, 0.05, 0.1))
        simple = float(np.sum(np.exp(data)))
    else:
        data = np.asarray(zscores)
        peak = _peak_fit(data, None, 0.3, (0.25, 2.0, 4.0))
        simple = float(np.sum(data))
    if peak is None:
        total = float(np.sum(data))
        if verbose:
            print(f"Fitting the double Gaussian failed. Return simple SUM = {total}")
        return 0.0, 0.0, 0.0, total
    above = data > peak
    total = simple if do_simple_sum > 0 else float(np.sum(above) * peak + np.sum(data[above]))
    if verbose:
        print("total_estimated_sum =", total)
    return 0.0, peak, 0.0, total


def score_index(results_file, do_run_CC, do_simple_sum, output_filename, verbose=True, base_dir=None):
    """Write the ReSuLt score file for every result file named in the index
    file ``results_file`` (one name per line; relative names resolve against
    ``base_dir`` when given, else the working directory)."""
    with open(results_file, "r") as fh:
        names = [ln.strip() for ln in fh.readlines()]
    with open(output_filename, "w") as out:
        for name in names:
            out.write(name + "\n")
            source = Path(name)
            if base_dir is not None and not source.is_absolute():
                source = Path(base_dir) / source
            _, cc_value, z_value, total = run_this(str(source), do_run_CC, do_simple_sum, verbose=verbose)
            out.write(f"{total} {z_value} {cc_value}\n")


def read_score(result_path) -> float:
    """The optimizer's objective from a ReSuLt file: minus the first number of
    its second line, NaN read as 0 (func.py:119-127)."""
    import math
    lines = Path(result_path).read_text().splitlines()
    if len(lines) < 2:
        raise RuntimeError(f"Incomplete score result: {result_path}")
    value = float(lines[1].split()[0])
    if math.isnan(value):
        value = 0.0
    if not math.isfinite(value):
        raise RuntimeError(f"Nonfinite score in {result_path}: {value}")
    return -value
