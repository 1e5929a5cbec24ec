"""The committed box-32 search fixture (tests/golden/box32_search.npz, written by
tests/golden/make_golden.py): the CPU oracle reproduces it (regression guard)
and the HIP path matches it (GPU)."""
import os
import sys

import numpy as np
import pytest

from oracle import search as osr
from oracle import splat as osp

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "box32_search.npz")


@pytest.fixture(scope="module")
def gold():
    return dict(np.load(G, allow_pickle=False))


def _params(g):
    box, apix, res, psi_step, local_range, crop, kk, mr, me = g["params"]
    return int(box), float(apix), float(res), float(psi_step), float(local_range), int(crop), float(kk), int(mr), int(me)


def test_oracle_reproduces_fixture(gold):
    box, apix, res, psi_step, local_range, crop, kk, mr, me = _params(gold)
    vol = osp.density_from_coords(gold["coords"], gold["numbers"], box, apix, res)
    np.testing.assert_array_equal(vol, gold["volume"])
    parts = {k: gold[k] for k in ("images", "rot", "tilt", "psi", "dfu", "dfv", "dfa")}
    recs = osr.run_search(vol.astype(np.float64), gold["dir_rot"], gold["dir_tilt"], parts, apix=apix,
                          new_box=box, psi_step=psi_step, local_range=local_range, crop=crop, kk=kk,
                          mask_radius=mr, mask_edge=me, dtype=np.float64)
    for k in ("line", "best_local", "best_psi", "xshift", "yshift"):
        assert [r[k] for r in recs] == list(gold["rec_" + k])
    np.testing.assert_allclose([r["rawcc"] for r in recs], gold["rec_rawcc"], rtol=1e-9)


@pytest.mark.gpu
def test_hip_matches_fixture(hip_device, gold):
    from gisapr_amd.engine import SearchEngine
    box, apix, res, psi_step, local_range, crop, kk, mr, me = _params(gold)
    print("[case] golden box32", flush=True)
    eng = SearchEngine(box=box, apix=apix, crop=crop, psi_step=psi_step, local_range=local_range, kk=kk,
                       mask_radius=mr, mask_edge=me, dir_rot=gold["dir_rot"], dir_tilt=gold["dir_tilt"],
                       device=hip_device)
    dfo = np.stack([gold["dfu"], gold["dfv"], gold["dfa"]], 1)
    ps = eng.add_particles(gold["images"], gold["rot"], gold["tilt"], gold["psi"], defocus=dfo)
    vol = eng.density(gold["coords"], gold["numbers"], res)
    res_ = eng.score_volume(vol, ps)
    recs = eng.records(res_, ps)
    for k in ("line", "best_local", "best_psi", "xshift", "yshift"):
        assert [r[k] for r in recs] == list(gold["rec_" + k]), k
    raw = np.asarray([r["rawcc"] for r in recs])
    np.testing.assert_allclose(raw, gold["rec_rawcc"], rtol=1e-4, atol=1e-3)
    assert abs(float(res_.score.cpu()) - float(gold["score"])) <= 1e-4 * float(gold["score"])
