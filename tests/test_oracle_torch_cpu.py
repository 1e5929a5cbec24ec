"""The torch-CPU restatement of the v606 path (oracle/torch_cpu.py, the CPU
baseline bench.py times) against the float64 numpy oracle on the same
inputs: template bank and particle spectra within 1e-4 relative, and the
per-particle peak score / best (direction, psi, shift) equal to the oracle's
search, except where two candidates are within float32 rounding (near-tie)."""
import numpy as np
import torch

from oracle import fourier as ofr
from oracle import particles as opt
from oracle import project as opj
from oracle import search as osr
from oracle import splat as osp
from oracle import torch_cpu as otc


def _setup(name="box32", B=12):
    from gisapr_amd import synthetic
    wl = synthetic.WORKLOADS[name]
    d = synthetic.make_workload(name, n_particles=B)
    vol = osp.density_from_coords(d["coords"], d["numbers"], wl.box, wl.apix, wl.res)
    N = wl.box
    mask = opj.generate_soft_mask((N, N), wl.mask_radius, wl.mask_edge)
    fsc = ofr.centered_full_to_packed(opj.expand_radial_values(np.ones(N, np.float32), N))
    return wl, d, vol, mask, fsc


def test_templates_and_particles_match_float64_oracle():
    wl, d, vol, mask, fsc = _setup()
    rot, tilt = d["dir_rot"][:20], d["dir_tilt"][:20]
    got = otc.build_templates(vol, rot, tilt, wl.box, mask, fsc).numpy()
    want = opj.build_templates(vol.astype(np.float64), rot, tilt, wl.box, mask, fsc, np.float64)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-4
    for i in range(len(rot)):  # every row, not only the bank as a whole
        assert np.linalg.norm(got[i] - want[i]) / np.linalg.norm(want[i]) < 1e-4, i
    ctf = opt.PackedCTF(wl.box, wl.apix, wl.cs, wl.voltage, wl.kk)
    img = d["images"][0].numpy()
    _, eff = opt.psi_sampling_for_record(float(d["psi"][0]), wl.psi_step, True, wl.local_range)
    c = ctf.one(d["dfu"][0], d["dfv"][0], d["dfa"][0])
    sp = otc.prepare_particle(img, eff, wl.box, mask, torch.from_numpy(c)).numpy()
    sw = opt.prepare_particle(img.astype(np.float64), eff, wl.box, mask, c.astype(np.float64), np.float64)
    assert np.linalg.norm(sp - sw) / np.linalg.norm(sw) < 1e-4


def test_search_matches_oracle_records():
    wl, d, vol, mask, fsc = _setup()
    parts = {"images": d["images"].numpy(), "rot": d["rot"], "tilt": d["tilt"], "psi": d["psi"],
             "dfu": d["dfu"], "dfv": d["dfv"], "dfa": d["dfa"]}
    recs, table = osr.run_search(vol.astype(np.float64), d["dir_rot"], d["dir_tilt"], parts, apix=wl.apix,
                                 new_box=wl.box, psi_step=wl.psi_step, local_range=wl.local_range, crop=wl.crop,
                                 kk=wl.kk, mask_radius=wl.mask_radius, mask_edge=wl.mask_edge, dtype=np.float64,
                                 return_scores=True)
    bank = otc.build_templates(vol, d["dir_rot"], d["dir_tilt"], wl.box, mask, fsc)
    local, eff = [], []
    for b in range(len(parts["images"])):
        local.append(osr.local_orientations(float(d["rot"][b]), float(d["tilt"][b]), d["dir_rot"], d["dir_tilt"],
                                            wl.local_range))
        eff.append(opt.psi_sampling_for_record(float(d["psi"][b]), wl.psi_step, True, wl.local_range)[1])
    ctf = opt.PackedCTF(wl.box, wl.apix, wl.cs, wl.voltage, wl.kk)
    dfo = np.stack([d["dfu"], d["dfv"], d["dfa"]], 1)
    got, nscores = otc.search(bank, parts["images"], dfo, eff, local, wl.box, mask, ctf, wl.crop, ctf_batch_size=5)
    assert nscores == sum(t.size for t in table)
    done = [g for g in got if g is not None]
    assert len(done) == len(recs) == len(table)
    for g, r, t in zip(done, recs, table):
        flat = t.reshape(-1)
        scale = np.abs(flat).max()
        assert abs(g[0] - r["rawcc"]) <= 1e-4 * scale
        if g[1] != int(np.argmax(flat)):
            assert flat.max() - flat[g[1]] <= 1e-4 * scale
