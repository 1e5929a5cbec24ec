"""GPU tests of the batched multi-candidate step (CandidateBatch): each slot's
score and per-particle peaks equal a single-candidate evaluation of the same
pose on the slot's particle subset, graph replays equal eager steps, and the
autotuner's variants give identical results.

Tolerance: the batched step splats the fixed body and the moving chain in two
passes (float atomics in a different order), so volumes agree to float32
rounding, and scores to 1e-5 relative."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def setup(hip_device):
    from gisapr_amd import synthetic
    from gisapr_amd.engine import SearchEngine
    wl = synthetic.WORKLOADS["box64"]
    data = synthetic.make_workload("box64")
    eng = SearchEngine(box=wl.box, apix=wl.apix, crop=wl.crop, psi_step=wl.psi_step, local_range=wl.local_range,
                       kk=wl.kk, mask_radius=wl.mask_radius, mask_edge=wl.mask_edge, fsc_values=data["fsc"],
                       dir_rot=data["dir_rot"], dir_tilt=data["dir_tilt"], device=hip_device)
    eng.enable_small_shift(resolution=8.0, half=3)
    dfo = np.stack([data["dfu"], data["dfv"], data["dfa"]], 1)
    ps = eng.add_particles(data["images"].numpy(), data["rot"], data["tilt"], data["psi"], defocus=dfo)
    eng.prepare_small_shift(ps)
    rng = np.random.default_rng(11)
    poses = np.concatenate([rng.normal(0, 2.0, (6, 3)), rng.normal(0, 0.7, (6, 3))], 1)
    return wl, data, eng, ps, poses


def single(eng, ps, wl, data, pose, path):
    from gisapr_amd.pdb_text import move_chain
    xyz = move_chain(data["coords"], data["chains"], wl.moving_chain, pose)
    vol = eng.density(xyz, data["numbers"], wl.res)
    bank = eng.templates(vol)
    if path == "mixed":
        res = eng.search_small_shift(bank, ps)
    elif path == "split":
        res = eng.search_split(bank, ps)
    else:
        res = eng.search(bank, ps)
    return res.corr.best.cpu().numpy(), res.corr.peak.cpu().numpy(), res.corr.scores.cpu().numpy().reshape(-1)


@pytest.mark.parametrize("path,group,operand", [("mixed", 8, "bf16"), ("fp32", 8, None), ("mixed", 2, "bf16"),
                                                ("fp32", 2, None), ("mixed", 8, "fp16"), ("split", 8, None),
                                                ("split", 2, None)])
def test_batch_slots_match_single_candidates(setup, path, group, operand):
    wl, data, eng, ps, poses = setup
    if operand == "fp16":
        eng.set_small_shift_operand("fp16")
    try:
        _batch_slots_match_single_candidates(setup, path, group, operand)
    finally:
        if operand == "fp16":
            eng.set_small_shift_operand("bf16")


def _batch_slots_match_single_candidates(setup, path, group, operand):
    from gisapr_amd.batch import CandidateBatch, RigidBody
    wl, data, eng, ps, poses = setup
    say("batch slots", path, "template group", group, "operand", operand)
    body = RigidBody.from_arrays(data["coords"], data["numbers"], data["chains"], wl.moving_chain)
    subsets = [np.arange(ps.B), np.arange(0, ps.B, 2), np.arange(5, 17)]
    batch = CandidateBatch(eng, ps, body, wl.res, slots=3, path=path, subsets=subsets, template_group=group)
    got = batch.evaluate(poses[:3])
    _, corr = batch.evaluate(poses[:3], sync=False)
    off = ps.dir_off.cpu().numpy()
    for c in range(3):
        best, peak, scores = single(eng, ps, wl, data, poses[c], path)
        sub = subsets[c]
        ok = best[sub] >= 0
        want = float(np.sum(np.exp(peak[sub][ok].astype(np.float64) / 1024.0)))
        assert got[c] == pytest.approx(want, rel=1e-5), (c, got[c], want)
        r = batch.slot_result(corr, c)
        np.testing.assert_allclose(r["peak"].cpu().numpy(), peak[sub], rtol=1e-5, atol=1e-3)
        # the same best pair for every particle, except a near-tie: the two
        # pairs' scores (single-candidate run) within 1e-5 of each other
        for k, b in enumerate(sub):
            bb = int(r["best"][k])
            if bb != best[b]:
                assert abs(scores[off[b] * ps.P + bb] - peak[b]) <= 1e-5 * abs(peak[b]), (c, b, bb, best[b])


@pytest.mark.parametrize("path", ["mixed", "fp32", "split"])
def test_batch_graph_replays_match_eager(setup, path):
    import torch
    from gisapr_amd.batch import CandidateBatch, RigidBody
    wl, data, eng, ps, poses = setup
    say("batch graph", path)
    body = RigidBody.from_arrays(data["coords"], data["numbers"], data["chains"], wl.moving_chain)
    eager = CandidateBatch(eng, ps, body, wl.res, slots=2, path=path)
    graphed = CandidateBatch(eng, ps, body, wl.res, slots=2, path=path)
    graphed.capture(poses[:2])
    for rep, (a, b) in enumerate([(0, 1), (2, 3), (4, 5), (0, 1)]):
        want = eager.evaluate(poses[[a, b]])
        x = torch.randn(1 << 20, device=eng.device)   # unrelated eager work between replays
        (x * 2).sum().item()
        got = graphed.evaluate(poses[[a, b]])
        torch.cuda.synchronize()
        np.testing.assert_allclose(got, want, rtol=1e-5, err_msg=f"replay {rep}")
    if path == "mixed":  # new lag tables (same values): the graph holds the old addresses and must be re-captured
        eng.set_small_shift_operand("bf16")
        with pytest.raises(RuntimeError, match="capture again"):
            graphed.evaluate(poses[:2])


def test_autotuner_variants_agree(setup, tmp_path):
    from gisapr_amd.batch import Autotuner, CandidateBatch, RigidBody, _variants_for
    wl, data, eng, ps, poses = setup
    say("autotuner")
    body = RigidBody.from_arrays(data["coords"], data["numbers"], data["chains"], wl.moving_chain)
    tuner = Autotuner(str(tmp_path / "tune.json"))
    for path in ("mixed", "fp32", "split"):
        batch = CandidateBatch(eng, ps, body, wl.res, slots=2, path=path, autotuner=tuner)
        tile = batch.tune(poses[:2])
        assert tile in _variants_for(path, ps.P)
        outs = []
        for v in _variants_for(path, ps.P):
            corr = batch._search(v)
            outs.append((corr.scores.cpu().numpy(), corr.best.cpu().numpy()))
        for s, b in outs[1:]:
            np.testing.assert_array_equal(s, outs[0][0])
            np.testing.assert_array_equal(b, outs[0][1])
    again = Autotuner(str(tmp_path / "tune.json"))
    assert again.table == tuner.table and len(tuner.table) == 3


def test_batch_search_engine_matches_run_search(hip_device, tmp_path):
    """ZL-shaped BatchSearchEngine.run: each model's records equal run_search
    on that model's density (same direction, psi and shift; rawCC within
    1e-5: the fixed body and the moving chain are splatted in two passes)."""
    import ast
    import torch
    from gisapr_amd import density, mrc, search, star, synthetic
    from gisapr_amd.batch import BatchSearchEngine
    say("BatchSearchEngine")
    wl = synthetic.WORKLOADS["box32"]
    d = synthetic.make_workload("box32")
    mrc.write_mrc(tmp_path / "parts.mrcs", d["images"].numpy())
    star.write_particle_star(tmp_path / "parts.star", str(tmp_path / "parts.mrcs"), d["rot"], d["tilt"], d["psi"],
                             d["dfu"], d["dfv"], d["dfa"])
    star.write_direction_star(tmp_path / "dirs.star", d["dir_rot"], d["dir_tilt"])
    (tmp_path / "model.pdb").write_text(synthetic.pdb_text(d["coords"], d["elements"], d["chains"]))
    argv = ["--i", "x", "--p", str(tmp_path / "parts.star"), "--o", str(tmp_path / "out.txt"),
            "--ang", str(tmp_path / "dirs.star"), "--mrc", "unused.mrc",
            "--oriboxsize", str(wl.box), "--newboxsize", str(wl.box),
            "--apix", str(wl.apix), "--kk", str(wl.kk), "--psiStep", str(wl.psi_step), "--doLocalSearch",
            "--localRange", str(wl.local_range), "--ccgCropSize", str(wl.crop),
            "--maskRadius", str(wl.mask_radius), "--maskEdge", str(wl.mask_edge)]
    args = search.create_search_parser().parse_args(argv)
    models = [dict(pdb_path=str(tmp_path / "model.pdb"), source_pdb=str(tmp_path / "model.pdb"),
                   chain_id=wl.moving_chain, pose=list(p), boxsize=wl.box, apix=wl.apix, res=wl.res)
              for p in ([0, 0, 0, 0, 0, 0], [3.0, -2.0, 1.0, 0.5, 0.0, -0.5], [-4.0, 1.0, 2.0, 0.0, 1.0, 0.0])]
    eng = BatchSearchEngine(None, hip_device, dict(batch_candidates=2))
    out = eng.run(args, models)
    assert len(out) == 3
    for m, r in zip(models, out):
        vol = density.generate_volume(**{k: v for k, v in m.items()}, device=hip_device)
        search.run_search(args, model_volume=vol)
        want = [ast.literal_eval(x) for x in open(tmp_path / "out.txt")]
        got = [ast.literal_eval(x) for x in r["search_text"].splitlines()]
        assert len(got) == len(want) > 0
        for g, w in zip(got, want):
            assert g[:2] == w[:2]
            if g[0] == "cc: ":
                assert g[5:7] == w[5:7] and g[8:10] == w[8:10]
            elif g[0] == "psi: ":
                assert g[2] == w[2]
            else:
                assert g[2] == pytest.approx(w[2], rel=1e-5)
        raw = [w[2] for w in want if w[0] == "rawCC: "]
        assert r["score"] == pytest.approx(float(np.sum(np.exp(np.asarray(raw) / 1024.0))), rel=1e-5)


def test_batch_search_engine_calibration_rows(hip_device, tmp_path):
    """BatchSearchEngine.run(calibrate=True) returns each model's Top-K rows;
    running a model restricted to its own rows keeps its best direction, psi,
    shift and rawCC; calibrating restricted directions is refused."""
    import ast
    from gisapr_amd import search, star, synthetic
    from gisapr_amd.batch import BatchSearchEngine
    say("BatchSearchEngine calibration")
    wl = synthetic.WORKLOADS["box32"]
    d = synthetic.make_workload("box32")
    from gisapr_amd import mrc
    mrc.write_mrc(tmp_path / "parts.mrcs", d["images"].numpy())
    star.write_particle_star(tmp_path / "parts.star", str(tmp_path / "parts.mrcs"), d["rot"], d["tilt"], d["psi"],
                             d["dfu"], d["dfv"], d["dfa"])
    star.write_direction_star(tmp_path / "dirs.star", d["dir_rot"], d["dir_tilt"])
    (tmp_path / "model.pdb").write_text(synthetic.pdb_text(d["coords"], d["elements"], d["chains"]))
    argv = ["--i", "x", "--p", str(tmp_path / "parts.star"), "--o", str(tmp_path / "out.txt"),
            "--ang", str(tmp_path / "dirs.star"), "--mrc", "unused.mrc",
            "--oriboxsize", str(wl.box), "--newboxsize", str(wl.box),
            "--apix", str(wl.apix), "--kk", str(wl.kk), "--psiStep", str(wl.psi_step), "--doLocalSearch",
            "--localRange", str(wl.local_range), "--ccgCropSize", str(wl.crop),
            "--maskRadius", str(wl.mask_radius), "--maskEdge", str(wl.mask_edge), "--keep-orientations", "2"]
    args = search.create_search_parser().parse_args(argv)
    models = [dict(pdb_path=str(tmp_path / "model.pdb"), source_pdb=str(tmp_path / "model.pdb"),
                   chain_id=wl.moving_chain, pose=list(p), boxsize=wl.box, apix=wl.apix, res=wl.res)
              for p in ([0, 0, 0, 0, 0, 0], [3.0, -2.0, 1.0, 0.5, 0.0, -0.5])]
    eng = BatchSearchEngine(None, hip_device, dict(batch_candidates=2))
    assert eng.options == dict(batch_candidates=2)  # kept as given: the worker compares it
    cal = eng.run(args, models, calibrate=True)
    assert eng.particles.bytes > 0 and eng.volumes.bytes == 0  # the worker's ping fields
    lines = {str(r.line_index) for r in star.read_particle_table(str(tmp_path / "parts.star")).records}
    for r in cal:
        assert set(r["orientation_rows"]) == lines
        assert all(len(row["directions"]) <= 2 for row in r["orientation_rows"].values())
    assert eng.run(args, models)[0]["orientation_rows"] is None
    for i in range(2):
        res = eng.run(args, [models[i]], orientation_rows=cal[i]["orientation_rows"])[0]
        got = [ast.literal_eval(x) for x in res["search_text"].splitlines()]
        want = [ast.literal_eval(x) for x in cal[i]["search_text"].splitlines()]
        assert len(got) == len(want) > 0
        for g, w in zip(got, want):
            assert g[:2] == w[:2]
            if g[0] == "cc: ":
                assert g[5:7] == w[5:7] and g[8:10] == w[8:10]
            elif g[0] == "psi: ":
                assert g[2:] == w[2:]
            else:  # the density's float atomics may add in another order between runs
                assert g[2] == pytest.approx(w[2], rel=1e-5)
    with pytest.raises(ValueError):
        eng.run(args, models[:1], orientation_rows=cal[0]["orientation_rows"], calibrate=True)
