"""Writes tests/golden/box32_search.npz: a seeded box-32 search (inputs and the
float64 oracle's per-particle records and score).

    python tests/golden/make_golden.py

The fixture is data only: the model atoms, direction table, particle images
and metadata, and the expected outputs.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from gisapr_amd import synthetic  # noqa: E402
from oracle import search as osr  # noqa: E402
from oracle import splat as osp  # noqa: E402


def main():
    wl = synthetic.WORKLOADS["box32"]
    d = synthetic.make_workload("box32")
    vol = osp.density_from_coords(d["coords"], d["numbers"], wl.box, wl.apix, wl.res)
    imgs = d["images"].numpy().astype(np.float32)
    parts = {"images": imgs, "rot": d["rot"], "tilt": d["tilt"], "psi": d["psi"],
             "dfu": d["dfu"], "dfv": d["dfv"], "dfa": d["dfa"]}
    recs = osr.run_search(vol.astype(np.float64), d["dir_rot"], d["dir_tilt"], parts, apix=wl.apix,
                          new_box=wl.box, psi_step=wl.psi_step, local_range=wl.local_range, crop=wl.crop,
                          kk=wl.kk, mask_radius=wl.mask_radius, mask_edge=wl.mask_edge, dtype=np.float64)
    keys = ["line", "best_local", "best_psi", "xshift", "yshift", "peak_y", "peak_x", "rawcc", "z", "mean", "std"]
    out = {f"rec_{k}": np.asarray([r[k] for r in recs]) for k in keys}
    np.savez_compressed(
        os.path.join(HERE, "box32_search.npz"),
        coords=d["coords"], numbers=d["numbers"], dir_rot=d["dir_rot"], dir_tilt=d["dir_tilt"],
        images=imgs, rot=d["rot"], tilt=d["tilt"], psi=d["psi"], dfu=d["dfu"], dfv=d["dfv"], dfa=d["dfa"],
        volume=vol, score=np.float64(osr.score_sum([r["rawcc"] for r in recs])),
        params=np.asarray([wl.box, wl.apix, wl.res, wl.psi_step, wl.local_range, wl.crop, wl.kk,
                           wl.mask_radius, wl.mask_edge], np.float64),
        **out)
    print("wrote", os.path.join(HERE, "box32_search.npz"), len(recs), "records")


if __name__ == "__main__":
    main()
