"""CPU checks of bench.py's contract: one GPU only, the work-per-score
figures behind `roofline`, and the committed traffic summaries it reads."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def test_split_search_work_per_score():
    # 3 bf16 products x (Re Q Re Phi - Im Q Im Phi: 2 multiplies + 2 adds) per kept coefficient and lag;
    # box 256 at 1.06 A/px inside 4 A keeps 7,288 coefficients, 49 lags
    assert bench.mixed_flops_per_score(7288, 49) == 3 * 4 * 7288 * 49 == 4285344


def test_traffic_summaries_are_read():
    for stage in ("mixed_search", "templates_stage", "fp32_search", "split_search", "split16_search"):
        t = bench.profiled_traffic(stage, "box256")
        assert t is not None and t > 1e8, stage
    assert bench.profiled_traffic("no_such_kernel", "box256") is None


def test_fp32_search_work_per_score():
    # the reference's full C2R per score: 6 n (n/2 + 1) for the product + 2.5 n^2 log2(n^2) for the transform
    assert bench.correlation_flops_per_score(256) == pytest.approx(6 * 256 * 129 + 2.5 * 256 ** 2 * 16)
    assert bench.correlation_flops_per_score(256) == pytest.approx(2.82e6, rel=1e-3)
