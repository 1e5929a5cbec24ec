"""The worker protocol on the GPU: every request through a persistent worker
process gives what the same call gives in process -- search_from_pdb
(density on the worker's GPU, then run_search; gisapr_worker.py:117-150),
search_batch_from_pdb (ZL gisapr_worker_v2.py:120-139), geometry
(func_check_boundary_for_testing_v8.py:62-126), and the served evaluate
requests (continuous batching) against one CandidateBatch per request.

The worker splats with float atomics, so volumes made in two processes may
differ in the last bits: scores are compared to 1e-5 relative, and the
per-particle records' rawCC to 1e-4 with their (rot, tilt, shifts) equal."""
import ast

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _records(path):
    recs = {}
    for line in open(path):
        item = ast.literal_eval(line)
        recs.setdefault(item[1], {})[item[0]] = item
    return recs


def _same_records(a, b):
    assert sorted(a) == sorted(b)
    for line in a:
        ca, cb = a[line]["cc: "], b[line]["cc: "]
        assert (ca[5], ca[6], ca[8], ca[9]) == (cb[5], cb[6], cb[8], cb[9]), line
        assert a[line]["rawCC: "][2] == pytest.approx(b[line]["rawCC: "][2], rel=1e-4, abs=1e-3)


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    from gisapr_amd import mrc, star, synthetic
    from oracle import splat as osp
    name = "box32"
    wl = synthetic.WORKLOADS[name]
    d = synthetic.make_workload(name)
    root = tmp_path_factory.mktemp("worker")
    mrc.write_mrc(root / "parts.mrcs", d["images"].numpy())
    star.write_particle_star(root / "parts.star", str(root / "parts.mrcs"), d["rot"], d["tilt"], d["psi"],
                             d["dfu"], d["dfv"], d["dfa"])
    star.write_direction_star(root / "dirs.star", d["dir_rot"], d["dir_tilt"])
    (root / "model.pdb").write_text(synthetic.pdb_text(d["coords"], d["elements"], d["chains"]))
    mrc.write_mrc(root / "unused.mrc", osp.density_from_coords(d["coords"][:4], d["numbers"][:4], wl.box, wl.apix,
                                                               wl.res))
    argv = ["--i", "x", "--p", str(root / "parts.star"), "--o", str(root / "out.txt"),
            "--ang", str(root / "dirs.star"), "--mrc", str(root / "unused.mrc"),
            "--oriboxsize", str(wl.box), "--newboxsize", str(wl.box), "--apix", str(wl.apix),
            "--kk", str(wl.kk), "--psiStep", str(wl.psi_step), "--doLocalSearch",
            "--localRange", str(wl.local_range), "--ccgCropSize", str(wl.crop),
            "--maskRadius", str(wl.mask_radius), "--maskEdge", str(wl.mask_edge)]
    return dict(root=root, wl=wl, argv=argv, pdb=str(root / "model.pdb"))


@pytest.fixture(scope="module")
def gpu_worker(hip_device):
    from gisapr_amd.runtime import PermanentWorker
    say("worker start")
    w = PermanentWorker(hip_device.index or 0, startup_timeout=300)
    yield w
    w.close()


def _with_out(argv, path):
    k = argv.index("--o")
    return argv[:k + 1] + [str(path)] + argv[k + 2:]


def test_search_from_pdb(hip_device, files, gpu_worker):
    from gisapr_amd import search
    from gisapr_amd.density import generate_volume
    wl, root = files["wl"], files["root"]
    model = dict(pdb_path=files["pdb"], boxsize=wl.box, apix=wl.apix, res=wl.res)
    say("search_from_pdb")
    got = gpu_worker.request("search_from_pdb", argv=_with_out(files["argv"], root / "w.txt"), model_kwargs=model)
    assert got["density_s"] > 0 and got["search_s"] > 0
    args = search.create_search_parser().parse_args(_with_out(files["argv"], root / "p.txt"))
    search.run_search(args, model_volume=generate_volume(**model, device=str(hip_device)))
    _same_records(_records(root / "w.txt"), _records(root / "p.txt"))


@pytest.mark.parametrize("path", ["split", "fp32"])
def test_search_batch_from_pdb(hip_device, files, gpu_worker, path):
    from gisapr_amd import search
    from gisapr_amd.batch import BatchSearchEngine
    wl = files["wl"]
    models = [dict(source_pdb=files["pdb"], chain_id=wl.moving_chain, pose=list(p), boxsize=wl.box, apix=wl.apix,
                   res=wl.res) for p in ([0, 0, 0, 0, 0, 0], [3.0, -2.0, 1.0, 0.5, 0.0, -0.5], [0, 0, 5.0, 1, 1, 0])]
    options = dict(batch_candidates=2, path=path)
    say("search_batch_from_pdb", path)
    got = gpu_worker.request("search_batch_from_pdb", argv=files["argv"], options=options, models=models)
    args = search.create_search_parser().parse_args(files["argv"])
    want = BatchSearchEngine(None, hip_device, options).run(args, models)
    assert len(got) == len(want) == 3
    for g, w in zip(got, want):
        assert g["score"] == pytest.approx(w["score"], rel=1e-5)
        ga = {ast.literal_eval(x)[1]: ast.literal_eval(x) for x in g["search_text"].splitlines()
              if x.startswith("['cc")}
        wa = {ast.literal_eval(x)[1]: ast.literal_eval(x) for x in w["search_text"].splitlines()
              if x.startswith("['cc")}
        assert {k: v[5:7] for k, v in ga.items()} == {k: v[5:7] for k, v in wa.items()}


def test_geometry(hip_device, files, gpu_worker):
    from gisapr_amd.restraint import GeometryRestraint
    wl = files["wl"]
    kwargs = dict(pdb_path=files["pdb"], chain_id=wl.moving_chain, boxsize=64, apix=1.5)
    g = GeometryRestraint(**kwargs, device=hip_device)
    for pose in ([0, 0, 0, 0, 0, 0], [0, 0, 0, 60.0, 0, 0]):
        say("geometry", pose)
        got = gpu_worker.request("geometry", geometry_kwargs=kwargs, pose=pose)
        assert tuple(got) == g.evaluate(pose)
    assert gpu_worker.request("ping")["geometry_cache_entries"] == 1


def test_served_evaluate_equals_alone(hip_device, files, gpu_worker):
    from gisapr_amd import search
    from gisapr_amd.batch import BatchSearchEngine, CandidateBatch, RigidBody
    wl = files["wl"]
    model = dict(source_pdb=files["pdb"], chain_id=wl.moving_chain, res=wl.res)
    say("add_dataset")
    # admission: a dataset that does not fit the HBM budget is refused, and the worker keeps serving
    with pytest.raises(Exception, match="MemoryError"):
        gpu_worker.request("add_dataset", name="ds32", argv=files["argv"], model_kwargs=model,
                           options=dict(path="split", slots=4, hbm_budget_mb=1))
    info = gpu_worker.request("add_dataset", name="ds32", argv=files["argv"], model_kwargs=model,
                              options=dict(path="split", slots=4))
    assert info["resident_bytes"] > 0 and info["particles"] > 0
    rng = np.random.default_rng(4)
    poses = [np.concatenate([rng.normal(0, 2.0, (k, 3)), rng.normal(0, 0.7, (k, 3))], 1) for k in (1, 5, 3, 2, 7)]
    say("evaluate", [len(p) for p in poses])
    rids = [gpu_worker.submit("ds32", p) for p in poses]
    replies = {r["rid"]: r for r in gpu_worker.collect(timeout=300)}
    assert sorted(replies) == sorted(rids)
    args = search.create_search_parser().parse_args(files["argv"])
    ctx = BatchSearchEngine(None, hip_device, dict(path="split"))._context(args)
    body = RigidBody.from_pdb(files["pdb"], wl.moving_chain)
    for rid, p in zip(rids, poses):
        r = replies[rid]
        assert r["ok"], r.get("error")
        alone = CandidateBatch(ctx["engine"], ctx["ps"], body, wl.res, slots=len(p), path="split").evaluate(p)
        np.testing.assert_allclose(r["result"]["scores"], alone, rtol=1e-5, err_msg=f"request {rid}")
        assert r["result"]["latency_s"] > 0
    with pytest.raises(Exception, match="already resident"):
        gpu_worker.request("add_dataset", name="ds32", argv=files["argv"], model_kwargs=model,
                           options=dict(path="split", slots=4))
    assert gpu_worker.request("ping")["datasets"] == ["ds32"]
