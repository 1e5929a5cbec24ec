"""Serving: the continuous-batching scheduler on the CPU, and on the GPU that
every request's scores are the same served in a mixed stream as evaluated
alone (within float32 atomic-order rounding, 1e-5 relative)."""
import sys

import numpy as np
import pytest

from gisapr_amd.serve import Request, Scheduler, latency_stats, poisson_requests


def test_scheduler_fifo_split_and_complete():
    s = Scheduler()
    a = Request(0, "d1", np.zeros((5, 6)), 0.0)
    b = Request(1, "d1", np.ones((2, 6)), 0.1)
    c = Request(2, "d2", np.ones((1, 6)), 0.2)
    for r in (a, b, c):
        s.arrive(r)
    assert s.pending("d1") == 7 and s.pending("d2") == 1
    step1 = s.take("d1", 4)
    assert [(r.rid, j) for r, j in step1] == [(0, 0), (0, 1), (0, 2), (0, 3)]
    step2 = s.take("d1", 4)                     # continuous batching: the tail of a joins b
    assert [(r.rid, j) for r, j in step2] == [(0, 4), (1, 0), (1, 1)]
    assert s.pending("d1") == 0
    s.finish(step2, [5.0, 6.0, 7.0], now=1.0)   # b completes before a
    assert [r.rid for r in s.completed] == [1] and b.t_done == 1.0
    assert np.allclose(b.scores, [6.0, 7.0])
    s.finish(step1, [1.0, 2.0, 3.0, 4.0], now=2.0)
    assert [r.rid for r in s.completed] == [1, 0]
    assert np.allclose(a.scores, [1, 2, 3, 4, 5]) and a.latency == 2.0
    assert not s.idle()
    s.finish(s.take("d2", 4), [9.0], now=3.0)
    assert s.idle()
    st = latency_stats(s.completed)
    assert st["n"] == 3 and st["p50_ms"] == pytest.approx(2000.0) and st["max_ms"] == pytest.approx(2800.0)


def test_poisson_requests_are_seeded_and_bounded():
    r1 = poisson_requests(300, 50.0, ["a", "b", "c"], seed=3)
    r2 = poisson_requests(300, 50.0, ["a", "b", "c"], seed=3)
    assert [(r.dataset, r.size, r.arrival) for r in r1] == [(r.dataset, r.size, r.arrival) for r in r2]
    sizes = np.array([r.size for r in r1])
    assert sizes.min() >= 1 and sizes.max() <= 64 and 5.0 < sizes.mean() < 11.0
    gaps = np.diff([0.0] + [r.arrival for r in r1])
    assert 0.012 < gaps.mean() < 0.03 and np.all(gaps > 0)
    with pytest.raises(ValueError):
        Request(9, "a", np.zeros((0, 6)), 0.0)


@pytest.mark.gpu
@pytest.mark.parametrize("path", ["mixed", "fp32", "split"])
def test_served_results_equal_alone(hip_device, path):
    from gisapr_amd import synthetic
    from gisapr_amd.batch import CandidateBatch, RigidBody
    from gisapr_amd.engine import SearchEngine
    from gisapr_amd.serve import SearchServer
    print(f"[case] serve mix vs alone ({path})", flush=True)
    server = SearchServer(hip_device, path=path, slots=4)
    parts = {}
    for name in ("box32", "box64"):
        wl = synthetic.WORKLOADS[name]
        data = synthetic.make_workload(name)
        eng = SearchEngine(box=wl.box, apix=wl.apix, crop=wl.crop, psi_step=wl.psi_step,
                           local_range=wl.local_range, kk=wl.kk, mask_radius=wl.mask_radius, mask_edge=wl.mask_edge,
                           fsc_values=data["fsc"], dir_rot=data["dir_rot"], dir_tilt=data["dir_tilt"],
                           device=hip_device)
        eng.enable_small_shift(resolution=wl.box * wl.apix / (0.3 * wl.box), half=3)
        dfo = np.stack([data["dfu"], data["dfv"], data["dfa"]], 1)
        ps = eng.add_particles(data["images"].numpy(), data["rot"], data["tilt"], data["psi"], defocus=dfo)
        body = RigidBody.from_arrays(data["coords"], data["numbers"], data["chains"], wl.moving_chain)
        server.add_dataset(name, eng, ps, body, wl.res)
        parts[name] = (eng, ps, body, wl.res)
    reqs = poisson_requests(14, 400.0, ["box32", "box64"], seed=5, mean_candidates=3, max_candidates=9)
    stats = server.serve(reqs)
    print("[case] stats", stats, flush=True)
    assert stats["requests"] == len(reqs) and stats["p99_ms"] >= stats["p50_ms"] > 0
    for r in reqs:
        eng, ps, body, res = parts[r.dataset]
        alone = CandidateBatch(eng, ps, body, res, slots=r.size, path=path).evaluate(r.poses)
        np.testing.assert_allclose(r.scores, alone, rtol=1e-5, err_msg=f"request {r.rid}")
    with pytest.raises(MemoryError):
        SearchServer(hip_device, hbm_budget_bytes=1 << 20, slots=2).add_dataset("box64", *parts["box64"])


def test_queue_growth_reports_a_growing_queue():
    from gisapr_amd.serve import queue_growth
    kept = [Request(i, "a", np.zeros((2, 6)), float(i)) for i in range(5)]
    for r in kept:
        r.t_done = r.arrival + 0.1  # constant latency: the server keeps up
    g = queue_growth(kept, 4.1)
    assert g["open_requests_at_last_arrival"] == 1 and abs(g["latency_growth_s_per_s"]) < 1e-12
    grow = [Request(i, "a", np.zeros((3, 6)), float(i)) for i in range(5)]
    for i, r in enumerate(grow):
        r.t_done = r.arrival + 0.5 * i + 0.1  # each request waits for all earlier ones
    g = queue_growth(grow, 6.1)
    assert g["open_requests_at_last_arrival"] == 2 and g["open_candidates_at_last_arrival"] == 6
    assert abs(g["latency_growth_s_per_s"] - 0.5) < 1e-9 and abs(g["drain_s"] - 2.1) < 1e-9
