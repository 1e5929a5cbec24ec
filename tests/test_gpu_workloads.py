"""GPU parity at the benchmark workloads' own sizes, against the float64 oracle.

Each case takes a seeded handful of particles of a workload (so the CPU
oracle finishes in seconds) and runs the oracle's search over the union of
their local directions (an order-preserving subset of the direction table,
so every particle's local list, and its local indices, are unchanged):
  * box 256 / 1.06 A (templates of transform length 512, particle spectra at
    256), FP32, split-bf16 and refined fp16 (split16) searches, crop 16;
  * the same particles at the reference's defaults newbox 160 and crop 32;
  * box 192 / 1.3 A (radix-3 transforms of length 384);
  * the batched step: two candidate poses on uneven particle subsets, each
    slot's score sum and peaks against the oracle on its moved model;
  * serving: one request per resident dataset (box 128 / 192 / 256), each
    candidate's score against the oracle;
  * on-grid noiseless box-96 particles: the search recovers the true
    direction and psi.
Tolerances: rawCC within 1e-4 relative and the same (direction, psi, shift)
for every particle; score sums within 1e-4 relative.
"""
import sys

import numpy as np
import pytest

from oracle import search as osr
from oracle import splat as osp

pytestmark = pytest.mark.gpu

RTOL = 1e-4


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def subset_data(name, n_particles):
    from gisapr_amd import synthetic
    wl = synthetic.WORKLOADS[name]
    data = synthetic.make_workload(name, n_particles=n_particles)
    return wl, data


def make_engine(wl, data, device, new_box=None, crop=None):
    from gisapr_amd.engine import SearchEngine
    return SearchEngine(box=wl.box, new_box=new_box, apix=wl.apix, crop=crop or wl.crop, psi_step=wl.psi_step,
                        local_range=wl.local_range, kk=wl.kk, voltage=wl.voltage, cs=wl.cs,
                        mask_radius=wl.mask_radius, mask_edge=wl.mask_edge, fsc_values=data["fsc"],
                        dir_rot=data["dir_rot"], dir_tilt=data["dir_tilt"], device=device)


def oracle_run(wl, data, coords, pick, new_box=None, crop=None):
    """Float64 oracle records of particles ``pick`` (dicts keyed by particle index) and their Σexp."""
    drot, dtilt = data["dir_rot"], data["dir_tilt"]
    union = sorted({i for b in pick for i in osr.local_orientations(data["rot"][b], data["tilt"][b], drot, dtilt,
                                                                     wl.local_range)})
    vol = osp.density_from_coords(coords, data["numbers"], wl.box, wl.apix, wl.res).astype(np.float64)
    imgs = data["images"].numpy() if hasattr(data["images"], "numpy") else np.asarray(data["images"])
    parts = {"images": imgs[pick], "rot": data["rot"][pick], "tilt": data["tilt"][pick], "psi": data["psi"][pick],
             "dfu": data["dfu"][pick], "dfv": data["dfv"][pick], "dfa": data["dfa"][pick]}
    recs = osr.run_search(vol, drot[union], dtilt[union], parts, apix=wl.apix, new_box=new_box or wl.box,
                          psi_step=wl.psi_step, local_range=wl.local_range, crop=crop or wl.crop, kk=wl.kk,
                          voltage=wl.voltage, cs=wl.cs, mask_radius=wl.mask_radius, mask_edge=wl.mask_edge,
                          fsc_values=data["fsc"], dtype=np.float64)
    out = {int(pick[r["line"]]): r for r in recs}
    return out, osr.score_sum([r["rawcc"] for r in recs])


def compare_records(got, want, tag):
    """got: engine records (line = position in the searched stack) keyed by particle."""
    assert sorted(got) == sorted(want), tag
    for b, w in want.items():
        g = got[b]
        assert (g["best_local"], g["best_psi"], g["xshift"], g["yshift"]) == \
               (w["best_local"], w["best_psi"], w["xshift"], w["yshift"]), (tag, b, g, w)
        assert abs(g["rawcc"] - w["rawcc"]) <= RTOL * abs(w["rawcc"]) + 1e-3, (tag, b, g["rawcc"], w["rawcc"])
        assert (g["rot"], g["tilt"]) == (w["rot"], w["tilt"]), (tag, b)


def engine_records(eng, data, pick, path, coords=None):
    imgs = data["images"].numpy()[pick]
    dfo = np.stack([data["dfu"], data["dfv"], data["dfa"]], 1)[pick]
    ps = eng.add_particles(imgs, data["rot"][pick], data["tilt"][pick], data["psi"][pick], defocus=dfo)
    vol = eng.density(data["coords"] if coords is None else coords, data["numbers"], data["workload"].res)
    bank, tinv = eng.templates(vol, normalize=False)
    if path == "split":
        res = eng.search_split(bank, ps, tinv)
    elif path == "split16":
        res = eng.search_split(bank, ps, tinv, operand="fp16")
    else:
        res = eng.search(bank, ps, tinv)
    recs = eng.records(res, ps)
    return {int(pick[r["line"]]): r for r in recs}, float(res.score.cpu())


@pytest.fixture(scope="module")
def box256(hip_device):
    wl, data = subset_data("box256", 8)
    pick = np.arange(8)
    say("box256 oracle (8 particles)")
    want, s_want = oracle_run(wl, data, data["coords"], pick)
    return wl, data, pick, want, s_want


@pytest.mark.parametrize("path", ["fp32", "split", "split16"])
def test_box256_search_matches_oracle(hip_device, box256, path):
    wl, data, pick, want, s_want = box256
    say("box256 search", path)
    got, s = engine_records(make_engine(wl, data, hip_device), data, pick, path)
    compare_records(got, want, f"box256 {path}")
    assert abs(s - s_want) <= RTOL * s_want


@pytest.mark.parametrize("path", ["fp32", "split", "split16"])
def test_reference_defaults_newbox160_crop32(hip_device, box256, path):
    """--newboxsize 160 (read_pdb_index_generate_sh_3DEG_local_v37.py:36) and
    --ccgCropSize 32 (v606:1562) on box-256 particles."""
    wl, data, pick, _, _ = box256
    pick = pick[:4]
    say("box256 newbox 160 crop 32", path)
    want, s_want = oracle_run(wl, data, data["coords"], pick, new_box=160, crop=32)
    got, s = engine_records(make_engine(wl, data, hip_device, new_box=160, crop=32), data, pick, path)
    compare_records(got, want, f"newbox160 {path}")
    assert abs(s - s_want) <= RTOL * s_want


def test_box192_search_matches_oracle(hip_device):
    wl, data = subset_data("box192", 6)
    pick = np.arange(6)
    say("box192 oracle (6 particles)")
    want, s_want = oracle_run(wl, data, data["coords"], pick)
    for path in ("fp32", "split", "split16"):
        say("box192 search", path)
        got, s = engine_records(make_engine(wl, data, hip_device), data, pick, path)
        compare_records(got, want, f"box192 {path}")
        assert abs(s - s_want) <= RTOL * s_want


def test_batched_step_matches_oracle(hip_device):
    """Two candidate poses of the box-192 moving chain on uneven particle
    subsets in one batched split step: per-slot Σexp and per-particle peaks
    against the oracle on each moved model (pdb_text.py:123-142 + v606)."""
    from gisapr_amd.batch import CandidateBatch, RigidBody
    from gisapr_amd.pdb_text import move_chain
    wl, data = subset_data("box192", 6)
    eng = make_engine(wl, data, hip_device)
    dfo = np.stack([data["dfu"], data["dfv"], data["dfa"]], 1)
    ps = eng.add_particles(data["images"].numpy(), data["rot"], data["tilt"], data["psi"], defocus=dfo)
    body = RigidBody.from_arrays(data["coords"], data["numbers"], data["chains"], wl.moving_chain)
    poses = np.array([[3.0, -2.0, 1.0, 0.8, -1.2, 0.5], [-1.5, 0.5, 4.0, -1.0, 0.3, 1.5]])
    subsets = [np.array([0, 2, 3, 5]), np.array([1, 4])]
    oracle = []
    for c in range(2):
        say("batched step oracle, slot", c)
        moved = move_chain(data["coords"], data["chains"], wl.moving_chain, poses[c])
        oracle.append(oracle_run(wl, data, moved, subsets[c]))
    for path in ("split16", "split", "fp32"):
        say("batched step vs oracle", path)
        batch = CandidateBatch(eng, ps, body, wl.res, slots=2, path=path, subsets=subsets)
        got = batch.evaluate(poses)
        _, corr = batch.evaluate(poses, sync=False)
        for c in range(2):
            want, s_want = oracle[c]
            assert abs(got[c] - s_want) <= RTOL * s_want, (path, c, got[c], s_want)
            r = batch.slot_result(corr, c)
            peaks = r["peak"].cpu().numpy()
            for k, b in enumerate(subsets[c]):
                assert abs(peaks[k] - want[int(b)]["rawcc"]) <= RTOL * abs(want[int(b)]["rawcc"]) + 1e-3


def test_serving_matches_oracle(hip_device):
    """One request per resident dataset (box 128 / 192 / 256, the serving
    workload's generators), served in one continuous-batching stream on the
    refined fp16 path; each candidate's score against the oracle on its moved model."""
    from gisapr_amd.batch import RigidBody
    from gisapr_amd.pdb_text import move_chain
    from gisapr_amd.serve import Request, SearchServer
    server = SearchServer(hip_device, path="split16", slots=2)
    sets = {}
    for name in ("serve128", "serve192", "serve256"):
        wl, data = subset_data(name, 3)
        eng = make_engine(wl, data, hip_device)
        dfo = np.stack([data["dfu"], data["dfv"], data["dfa"]], 1)
        ps = eng.add_particles(data["images"].numpy(), data["rot"], data["tilt"], data["psi"], defocus=dfo)
        server.add_dataset(name, eng, ps, RigidBody.from_arrays(data["coords"], data["numbers"], data["chains"],
                                                              wl.moving_chain), wl.res)
        sets[name] = (wl, data)
    rng = np.random.default_rng(2)
    reqs = [Request(i, name, np.concatenate([rng.normal(0, 2.0, (1, 3)), rng.normal(0, 1.0, (1, 3))], 1),
                    0.002 * i) for i, name in enumerate(sets)]
    say("serving three datasets")
    stats = server.serve(reqs)
    assert stats["requests"] == 3
    for r in reqs:
        wl, data = sets[r.dataset]
        moved = move_chain(data["coords"], data["chains"], wl.moving_chain, r.poses[0])
        say("serving oracle", r.dataset)
        _, s_want = oracle_run(wl, data, moved, np.arange(3))
        assert abs(r.scores[0] - s_want) <= RTOL * s_want, (r.dataset, r.scores, s_want)


def test_search_recovers_on_grid_truth_box96(hip_device):
    """Noiseless, CTF-free box-96 particles projected from the atoms in real
    space at grid directions and grid psi (integer shifts): the HIP search
    with the truth as prior picks the true direction and psi."""
    import torch
    from gisapr_amd import synthetic
    wl = synthetic.WORKLOADS["box96"]
    coords, elements, _ = synthetic.random_walk_chains(wl.chain_sizes, confine=wl.confine)
    numbers = np.asarray([synthetic.ELEMENT_Z.get(e, 0.0) for e in elements])
    drot, dtilt = synthetic.healpix_nest_directions(wl.healpix_order)
    rng = np.random.default_rng(21)
    B = 32
    k = rng.integers(0, len(drot), B)
    psi = wl.psi_step * rng.integers(1, 47, B)
    shifts = rng.integers(-3, 4, (B, 2))
    imgs = synthetic.whiten_batch(synthetic.project_atoms(coords, numbers, drot[k], dtilt[k], psi, shifts,
                                                          wl.box, wl.apix, wl.res)).numpy()
    from gisapr_amd.engine import SearchEngine
    eng = SearchEngine(box=wl.box, apix=wl.apix, crop=wl.crop, psi_step=wl.psi_step, local_range=wl.local_range,
                       kk=wl.kk, mask_radius=wl.mask_radius, mask_edge=wl.mask_edge, dir_rot=drot, dir_tilt=dtilt,
                       device=hip_device)
    ps = eng.add_particles(imgs, drot[k], dtilt[k], psi)
    vol = eng.density(coords, numbers, wl.res)
    for path in ("fp32", "split", "split16"):
        bank, tinv = eng.templates(vol, normalize=False)
        if path == "fp32":
            res = eng.search(bank, ps, tinv)
        else:
            res = eng.search_split(bank, ps, tinv, operand="fp16" if path == "split16" else "bf16")
        recs = eng.records(res, ps)
        hit_dir = sum(r["direction"] == k[r["line"]] for r in recs)
        hit_psi = sum(r["psi_fake"] == psi[r["line"]] for r in recs)
        say("truth box96", path, hit_dir, hit_psi, len(recs))
        assert len(recs) == B and hit_dir >= B - 2 and hit_psi >= B - 2, (path, hit_dir, hit_psi)
