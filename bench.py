#!/usr/bin/env python
"""Candidate-scoring throughput on one MI355X.

One step = one candidate evaluation on the box-256 workload (12,000-atom
3-chain model, moving chain 3,000 atoms, 5,000 particles of 256x256, HEALPix
order 3 local search +/-10 deg, psi step 3.75 deg, crop 16):
move the chain (host), splat the full model, 3-D transform, 768-direction
template bank, prepare all particle spectra (rotation, FFT, CTF), correlate
every (particle, local direction, psi), per-particle reduction, and the
float64 score sum.  Inputs (particle images, metadata) are resident in HBM
before timing starts.

Prints ONE JSON line.  ``--gpus`` above 1 is not supported (one-GPU build).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

METRIC = "particle-orientation scores/s (mixed vs FP32 FFT, same top-1); p50/p99 ms"
PEAK_FP32_MFMA_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def correlation_flops_per_score(n: int) -> float:
    """Algorithmic FP32 work of one score: conj-multiply + full C2R."""
    return 6.0 * n * (n / 2 + 1) + 2.5 * n * n * math.log2(n * n)


def cpu_baseline(eng, ps, bank, sample: int):
    """The oracle (float32 NumPy port, one core) on the first ``sample`` particles:
    particle preparation + correlation + per-particle record, templates given."""
    from oracle import fourier as ofr
    from oracle import particles as opt
    from oracle import search as osr
    bank_pre = ofr.pre_shift(bank.cpu().numpy())
    mask = eng.mask.cpu().numpy() if eng.mask is not None else None
    ctf = opt.PackedCTF(eng.n, eng.new_apix, 2.7, 300.0, eng.ctf.kk)
    imgs = ps.images[:sample].cpu().numpy()
    dfo = ps.defocus[:sample].cpu().numpy()
    scores = 0
    t0 = time.perf_counter()
    for b in range(min(sample, ps.B)):
        idx = ps.local[b]
        if not idx:
            continue
        c = ctf.one(*dfo[b])
        pf = opt.prepare_particle(imgs[b], list(ps.eff_psi[b]), eng.n, mask, c, np.float32)
        sc, ys, xs = ofr.correlate(bank_pre[idx], pf, eng.crop, np.float32)
        osr.make_record(b, (ps.rot[b], ps.tilt[b], ps.psi[b]), list(ps.fake_psi[b]), eng.dir_rot,
                        eng.dir_tilt, idx, sc, ys, xs, eng.crop, eng.trans_scale)
        scores += sc.size
    dt = time.perf_counter() - t0
    return {"value": scores / dt, "unit": "scores/s", "cores": 1, "kind": "port",
            "sample": f"oracle (numpy float32, 1 thread) particle prep + correlation + record for the "
                      f"first {sample} particles of the same workload, templates from the GPU bank: "
                      f"{scores} scores in {dt:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="box256")
    ap.add_argument("--particles", type=int, default=None)
    ap.add_argument("--cpu-sample", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1", file=sys.stderr)
        return 2

    import torch
    from gisapr_amd import synthetic
    from gisapr_amd.engine import SearchEngine
    from gisapr_amd.pdb_text import move_chain

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    wl = synthetic.WORKLOADS[args.workload]
    t_setup = time.perf_counter()
    data = synthetic.make_workload(args.workload, n_particles=args.particles, device=dev)
    eng = SearchEngine(box=wl.box, apix=wl.apix, crop=wl.crop, psi_step=wl.psi_step,
                       local_range=wl.local_range, kk=wl.kk, voltage=wl.voltage, cs=wl.cs,
                       mask_radius=wl.mask_radius, mask_edge=wl.mask_edge, fsc_values=data["fsc"],
                       dir_rot=data["dir_rot"], dir_tilt=data["dir_tilt"], device=dev)
    dfo = np.stack([data["dfu"], data["dfv"], data["dfa"]], 1)
    ps = eng.add_particles(data["images"], data["rot"], data["tilt"], data["psi"], defocus=dfo, prepare=False)
    del data["images"]
    n_scores = ps.n_scores
    rng = np.random.default_rng(synthetic.SEED_CAND)
    poses = [synthetic.candidate_pose(seed=int(s), rot_deg=wl.cand_rot_deg, trans_A=wl.cand_trans_A)
             for s in rng.integers(0, 2 ** 31, args.warmup + args.steps)]
    coords, numbers, chains = data["coords"], data["numbers"], data["chains"]
    setup_s = time.perf_counter() - t_setup

    stages = ("density", "templates", "particles", "correlate")
    ev_time = {k: 0.0 for k in stages}

    def step(pose, timed):
        xyz = move_chain(coords, chains, wl.moving_chain, pose)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)] if timed else None
        if timed:
            ev[0].record()
        vol = eng.density(xyz, numbers, wl.res)
        if timed:
            ev[1].record()
        bank = eng.templates(vol)
        if timed:
            ev[2].record()
        eng.prepare(ps)
        if timed:
            ev[3].record()
        res = eng.search(bank, ps)
        if timed:
            ev[4].record()
        return res, bank, ev

    for i in range(args.warmup):
        step(poses[i], False)
    torch.cuda.synchronize()
    evs = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        res, bank, ev = step(poses[args.warmup + i], True)
        evs.append(ev)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    for ev in evs:
        for k, name in enumerate(stages):
            ev_time[name] += ev[k].elapsed_time(ev[k + 1])
    stage_ms = {k: v / args.steps for k, v in ev_time.items()}
    ms_per_step = dt * 1000.0 / args.steps
    value = n_scores * args.steps / dt
    score = float(res.score.cpu())

    corr_ms = stage_ms["correlate"]
    flops = correlation_flops_per_score(eng.n) * n_scores
    achieved = flops / (corr_ms * 1e-3) / 1e12
    out = {
        "metric": METRIC, "value": value, "unit": "scores/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic (seeded generators; random-walk model)",
        "config": {"workload": f"{wl.name}: {sum(wl.chain_sizes)}-atom model (moving chain "
                               f"{wl.chain_sizes[-1]}), {ps.B} particles {wl.box}x{wl.box} at {wl.apix} A/px, "
                               f"HEALPix order {wl.healpix_order} ({len(eng.dir_rot)} directions) local "
                               f"+/-{wl.local_range} deg, psi step {wl.psi_step} ({ps.P} psi), crop {wl.crop}, "
                               f"one candidate per step",
                   "box": wl.box, "particles": ps.B, "directions": len(eng.dir_rot), "psi": ps.P,
                   "crop": wl.crop, "scores_per_step": n_scores, "path": "fp32-fft"},
        "candidates_per_s": 1000.0 / ms_per_step,
        "stage_ms": stage_ms,
        "score": score,
        "setup_s": setup_s,
        "roofline": {"bound": "mfma", "kernel": "correlate_kernel", "achieved": achieved,
                     "peak": PEAK_FP32_MFMA_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_MFMA_TFLOPS,
                     "traffic": None,
                     "flops_per_score": correlation_flops_per_score(eng.n), "launch_ms": corr_ms},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline and args.cpu_sample > 0:
        out["cpu_baseline"] = cpu_baseline(eng, ps, bank, args.cpu_sample)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
