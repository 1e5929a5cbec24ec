#!/usr/bin/env python
"""Candidate-scoring throughput and serving latency on one MI355X.

Main workload (box256): 12,000-atom 3-chain model, moving chain 3,000 atoms,
5,000 particles of 256x256 at 1.06 A/px, HEALPix order 3 local search
+/-10 deg, psi step 3.75 deg (5 psi).  The particle stack is resident in HBM
and its psi spectra are prepared once per dataset before timing (reported
as ``particle_prep_ms``).  A candidate evaluation = move the chain (host),
splat, 3-D transform, 768-direction template bank, search of every
(particle, local direction, psi) pair, per-particle reduction and the
float64 score sum.  Legs, on the same particles and poses:

  fp32           one candidate per step, the v606 FP32 search (every packed
                 coefficient, the 16 x 16 crop: +/-8 px) on FP32 MFMA
  fft32          the same search on the fp32 leg's last bank, computed the
                 way the reference computes it (per pair a float32 irfft2 of
                 the product, torch / rocFFT on the GPU); search timed alone
  split          one candidate per step, the same search (same operator,
                 same crop) on split-bf16 MFMA
  split16        one candidate per step, the same search with fp16 operands
                 (packed fp16 products, FP32 accumulation), then each
                 particle's best pair and its near ties rescored on
                 split-bf16 MFMA: the split leg's answers, bit for bit
  mixed          one candidate per step, the small-shift split-bf16 search of
                 the precision variant (coefficients inside 1/(4 A), 49 lags,
                 +/-3 px): a different, truncated functional
  batched_fp32     ``--slots`` candidates per step, one search launch over
  batched_split    all of them, replayed as a captured HIP graph, tile
  batched_split16  variant from the autotuner (batched_mixed only when
  batched_mixed    asked for)

``value`` is batched_split16 particle-orientation scores/s (batched_split
when that leg is not run); beside it: the ratio to the plain fp32 leg, top-1
(direction, psi, shift) agreement of the split / split16 and fp32 searches
on every particle, the split searches' error against the same operator in
float64 on a 256-particle subset, and the same for the small-shift variant
(whose top-1 is not expected to match: it truncates the coefficients and the
lags).

Batched PSO leg (``--pso-iters``, 0 to skip): box-192 workload (20,000-atom
model, moving chain 4,500, 20,000 particles at 1.3 A/px), a swarm of 64
candidates x 20 iterations with seeded poses, each member scoring its own
2,000-20,000-particle subset, one graph-captured batched step per iteration.

Serving legs (``--serve-requests``, 0 to skip): three resident datasets (box
128 / 2,000, box 192 / 10,000, box 256 / 20,000 particles, 3,000-atom moving
chain), a seeded Poisson stream of requests at ``--serve-rate``/s with 1-64
candidates each (geometric, mean 8), continuous batching; p50/p99 request
latency and throughput on the refined fp16 path (``serving``: the full
2,000-request stream, the top-level p50/p99) and on the FP32 path
(``serving_fp32``, ``--serve-requests-fp32`` requests: the FP32 search is
~3x slower, so its stream is shorter).

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
PEAK_BF16_MFMA_TFLOPS = 2500.0
PEAK_HBM_GBS = 8000.0


def correlation_flops_per_score(n: int) -> float:
    """Algorithmic FP32 work of one v606 score: conj-multiply + full C2R."""
    return 6.0 * n * (n / 2 + 1) + 2.5 * n * n * math.log2(n * n)


def split_flops_per_score(n: int, crop: int) -> float:
    """Split-bf16 work of one full-plane score: stage 1 (columns to lags,
    complex, 8 flops per MAC) over n x (n/2+1) coefficients x C lags, stage 2
    (rows to lags, real part, 4 flops) over n rows x C^2 lags, 3 products each."""
    return 3.0 * (8.0 * n * (n // 2 + 1) * crop + 4.0 * n * crop * crop)


def split16_flops_per_score(n: int, crop: int) -> float:
    """MFMA work of one fp16-operand score: stage 1 with one fp16 product
    (8 flops per complex MAC) over n x (n/2+1) coefficients x C lags, stage 2
    on split-bf16 (3 products) over n rows x C^2 lags.  The split-bf16
    rescoring of each particle's best pair and near ties is extra work, not
    counted."""
    return 8.0 * n * (n // 2 + 1) * crop + 3.0 * 4.0 * n * crop * crop


def template_bytes_per_direction(N: int, n: int) -> float:
    """HBM bytes of one template row: the packed row written plus the
    trilinear slice gather from the padded 3-D half spectrum (8 taps of 8
    bytes per sample over 2N x (N+1) samples, uncached worst case)."""
    return 8.0 * n * (n // 2 + 1) + 8.0 * 8.0 * 2 * N * (N + 1)


def mixed_flops_per_score(K: int, lags: int) -> float:
    """Split-bf16 work of one small-shift score: 3 products x 4 flops per
    complex coefficient and lag (real part only)."""
    return 3.0 * 4.0 * K * lags


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def _median_time(fn, runs: int):
    times, out = [], None
    for _ in range(runs):
        t0 = time.perf_counter()
        out = fn()
        times.append(time.perf_counter() - t0)
    return float(np.median(times)), out


def cpu_baseline(gpu_scores_per_candidate: int, box256_particles: int, runs: int = 5, template_dirs: int = 64,
                 log=print):
    """The v606 FP32 path restated in torch on the host CPU (oracle.torch_cpu,
    the reference's chunking: 16 templates x 32 psi, CTF batches of 256,
    8 projections per batch), timed at the default thread count and at the
    reference default of 4 threads, median of ``runs`` runs after one
    warm-up:
      box96  (the CPU-runnable workload, 200 particles): search only, and
             density + templates (all 192 directions) + search;
      box256 (``box256_particles`` particles of the 5,000): search only; the
             volume transform timed, the per-direction template work timed
             on ``template_dirs`` of the 768 directions and scaled to 768
             (density and templates: median of 3); the 5,000-particle
             candidate extrapolated
             from the per-score search time and the GPU run's score count.
    ``value``: the extrapolated box256 candidate (density + 768 templates +
    every score of the 5,000 particles) in scores/s at the default thread
    count."""
    import torch
    from gisapr_amd import synthetic
    from oracle import fourier as ofr
    from oracle import particles as opt
    from oracle import project as opj
    from oracle import search as osr
    from oracle import splat as osp
    from oracle import torch_cpu as otc

    def inputs(name, B):
        wl = synthetic.WORKLOADS[name]
        d = synthetic.make_workload(name, n_particles=B)
        N = wl.box
        mask = opj.generate_soft_mask((N, N), wl.mask_radius, wl.mask_edge)
        fscv = d.get("fsc")
        fscv = np.ones(N, np.float32) if fscv is None else np.asarray(fscv, np.float32)
        fsc = ofr.centered_full_to_packed(opj.expand_radial_values(fscv, N))
        local = [osr.local_orientations(float(d["rot"][b]), float(d["tilt"][b]), d["dir_rot"], d["dir_tilt"],
                                        wl.local_range) for b in range(B)]
        eff = [opt.psi_sampling_for_record(float(d["psi"][b]), wl.psi_step, True, wl.local_range)[1]
               for b in range(B)]
        ctf = opt.PackedCTF(N, wl.apix, wl.cs, wl.voltage, wl.kk)
        dfo = np.stack([d["dfu"], d["dfv"], d["dfa"]], 1)
        # the searched bank: the directions some particle searches (rows are
        # independent of each other), local lists renumbered into it
        used = sorted({i for l in local for i in l})
        where = {i: k for k, i in enumerate(used)}
        vol = osp.density_from_coords(d["coords"], d["numbers"], N, wl.apix, wl.res)
        bank = otc.build_templates(vol, [d["dir_rot"][i] for i in used], [d["dir_tilt"][i] for i in used], N, mask,
                                   fsc)
        local = [[where[i] for i in l] for l in local]
        return wl, d, mask, fsc, local, eff, ctf, dfo, d["images"].numpy(), vol, bank

    default_threads = torch.get_num_threads()
    results = {}
    w96 = inputs("box96", synthetic.WORKLOADS["box96"].n_particles)
    w256 = inputs("box256", int(box256_particles))
    for threads in sorted({default_threads, 4}, reverse=True):
        torch.set_num_threads(threads)
        row = {}
        for key, (wl, d, mask, fsc, local, eff, ctf, dfo, imgs, vol, bank) in (("box96", w96), ("box256", w256)):
            N = wl.box
            M = len(d["dir_rot"])
            dirs = M if key == "box96" else min(int(template_dirs), M)

            def density():
                return osp.density_from_coords(d["coords"], d["numbers"], N, wl.apix, wl.res)

            def spectrum():
                return otc.volume_dft(vol)

            dft = spectrum()

            def templates():
                return otc.build_templates(vol, d["dir_rot"][:dirs], d["dir_tilt"][:dirs], N, mask, fsc, dft=dft)

            def search():
                return otc.search(bank, imgs, dfo, eff, local, N, mask, ctf, wl.crop)

            search()  # warm-up (FFT plans)
            t_search, (_, n_sc) = _median_time(search, runs)
            few = runs if key == "box96" else min(runs, 3)
            t_density, _ = _median_time(density, few)
            t_dft, _ = _median_time(spectrum, few)
            t_tpl, _ = _median_time(templates, few)
            t_tpl_all = t_dft + t_tpl * M / dirs  # one volume transform, then the per-direction work
            r = {"particles": len(imgs), "scores": int(n_sc), "directions": M, "search_s": t_search,
                 "search_scores_per_s": n_sc / t_search, "density_s": t_density, "templates_s": t_tpl_all}
            if dirs != M:
                r["templates_timed_directions"] = dirs
            if key == "box96":
                total = t_density + t_tpl_all + t_search
                r.update(candidate_s=total, candidates_per_s=1.0 / total, end_to_end_scores_per_s=n_sc / total)
            else:
                per_score = t_search / n_sc
                total = t_density + t_tpl_all + per_score * gpu_scores_per_candidate
                r.update(extrapolated_scores=int(gpu_scores_per_candidate), candidate_s_extrapolated=total,
                         candidates_per_s_extrapolated=1.0 / total,
                         end_to_end_scores_per_s_extrapolated=gpu_scores_per_candidate / total)
            row[key] = r
            log("cpu baseline", threads, key, r)
        results[str(threads)] = row
    torch.set_num_threads(default_threads)
    head = results[str(default_threads)]["box256"]
    return {"value": head["end_to_end_scores_per_s_extrapolated"], "unit": "scores/s", "cores": default_threads,
            "kind": "port", "cpu_model": _cpu_model(), "os_cpu_count": os.cpu_count(), "runs": runs,
            "sample": f"oracle.torch_cpu (torch {torch.__version__} float32 on the host; v606 chunking) -- value: "
                      f"box256 candidate extrapolated to {gpu_scores_per_candidate} scores (5,000 particles) "
                      f"from {box256_particles} timed particles, density + 768 templates "
                      f"(timed on {min(int(template_dirs), 768)}) + search, {default_threads} threads; "
                      f"median of {runs}",
            "threads": results}


def top1(res, ps, center):
    """Per particle (best flat pair, y shift, x shift); -1 pair for none."""
    best = res.corr.best.cpu().numpy()
    py = res.corr.peak_y.cpu().numpy().reshape(-1)
    px = res.corr.peak_x.cpu().numpy().reshape(-1)
    off = ps.dir_off.cpu().numpy()
    out = np.full((ps.B, 3), -1, np.int64)
    for b in range(ps.B):
        if best[b] < 0:
            continue
        g = off[b] * ps.P + best[b]
        out[b] = (best[b], center - py[g], center - px[g])
    return out


def fft32_search(eng, ps, bank, tinv, chunk: int = 1024):
    """The v606 search computed as the reference computes it (v606:1418-1449),
    in torch on the GPU: per (particle, local direction, psi) pair the product
    of the pre-shifted template and particle spectra (v606:535-537), one
    float32 ``torch.fft.irfft2`` (rocFFT), the C x C crop of the unshifted
    correlation and its maximum times 1024 n^2; pairs in chunks of ``chunk``.
    Returns (scores [pairs], y, x, best [B], peak [B]) in the layout of the
    searches' results (first maximum per particle in (direction, psi) order).
    A measurement baseline, not a product path."""
    import torch
    from gisapr_amd.search import PackedFourierLayout
    n, C, P, dev = eng.n, eng.crop, ps.P, eng.device
    lay = PackedFourierLayout(n, C, dev)
    Ts = torch.fft.fftshift(bank.unpack() * tinv[:, None, None], dim=(-2, -1))
    Ps = torch.fft.fftshift(eng.row_spectra(ps) * ps.inv_norm[..., None, None], dim=(-2, -1)).reshape(-1, n, n // 2 + 1)
    rows = ps.dir_idx.long()
    counts = (ps.dir_off[1:] - ps.dir_off[:-1]).long()
    row_particle = torch.repeat_interleave(torch.arange(ps.B, device=dev), counts)
    npairs = int(rows.numel()) * P
    scores = torch.empty(npairs, dtype=torch.float32, device=dev)
    index = torch.empty(npairs, dtype=torch.long, device=dev)
    mult = float(1024 * n * n)
    for k0 in range(0, npairs, chunk):
        k = torch.arange(k0, min(k0 + chunk, npairs), device=dev)
        r, p = k // P, k % P
        prod = Ts.index_select(0, rows[r]) * Ps.index_select(0, row_particle[r] * P + p)
        ccg = torch.fft.irfft2(prod, s=(n, n), dim=(-2, -1))
        flat = lay.centered_ccg_crop_from_unshifted(ccg).reshape(len(k), -1)
        v, i = torch.max(flat, dim=-1)
        scores[k0:k0 + len(k)] = v * mult
        index[k0:k0 + len(k)] = i
    seg = torch.repeat_interleave(row_particle, P)
    peak = torch.full((ps.B,), -float("inf"), device=dev).scatter_reduce(0, seg, scores, "amax")
    pos = torch.arange(npairs, device=dev)
    first = torch.where(scores == peak[seg], pos, torch.full_like(pos, npairs))
    gbest = torch.full((ps.B,), npairs, dtype=torch.long, device=dev).scatter_reduce(0, seg, first, "amin")
    best = torch.where(counts > 0, gbest - ps.dir_off[:-1].long() * P, torch.full_like(gbest, -1))
    return scores, index // C, index % C, best, peak


def precision_check(eng, ps, tk, res, subset: int, seed: int):
    """Mixed-leg scores against the same truncated functional in float64
    (torch complex128 on the device) on a seeded particle subset.  ``tk`` are
    the template coefficients of the candidate the mixed leg searched last."""
    import torch
    tab = eng.small
    dev = eng.device
    rng = np.random.default_rng(seed)
    pick = np.sort(rng.choice(ps.B, min(subset, ps.B), replace=False))
    tk = tk[:, :tab.K].to(torch.complex128)
    pk = ps.coeffs[:, :tab.K].to(torch.complex128)
    phi = torch.from_numpy(tab.phi).to(dev)
    off = ps.dir_off.cpu().numpy()
    idx = ps.dir_idx.cpu().numpy()
    sc = res.corr.scores.cpu().numpy().reshape(-1)
    best = res.corr.best.cpu().numpy()
    P = ps.P
    agree, worst, used = 0, 0.0, 0
    for b in pick:
        d = idx[off[b]:off[b + 1]]
        if len(d) == 0:
            continue
        q = (tk[torch.from_numpy(d).to(dev)][:, None, :] * pk[b * P:(b + 1) * P][None]).reshape(-1, tab.K)
        s = 1024.0 * (q @ phi).real                     # [pairs, lags]
        v = s.max(dim=1).values.cpu().numpy()
        got = sc[off[b] * P:off[b + 1] * P].astype(np.float64)
        worst = max(worst, float(np.abs(got - v).max() / np.abs(v).max()))
        agree += int(int(np.argmax(v)) == int(best[b]))
        used += 1
    return {"subset": int(used), "seed": seed, "top1_agree_fp64": agree / max(used, 1),
            "max_rel_err_fp64": worst, "reference": "same functional in float64 (torch complex128)"}


def precision_check_split(eng, ps, bank, res, subset: int, seed: int):
    """Split-search scores against the same operator in float64 (torch
    complex128 on the device, the separable functional W^T Q E of
    ``host.crop_functional``) on a seeded particle subset; ``bank`` =
    (unnormalised rows, 1/norms) of the candidate the split leg searched last.
    A top-1 difference counts as a near-tie when the float64 scores of the two
    choices differ by at most 1e-5 of the particle's best score.
    ``max_rel_err_fp64``: the particle's reported best score (rawCC) against
    the float64 best; ``max_rel_err_fp64_all_pairs``: every pair's score
    against the float64 one, relative to the particle's best."""
    import torch
    from gisapr_amd import host, kernels
    dev = eng.device
    raw, tinv = bank
    W, E = host.crop_functional(eng.n, eng.crop)
    Wt = torch.from_numpy(W.T.copy()).to(dev)
    Et = torch.from_numpy(E).to(dev)
    rng = np.random.default_rng(seed)
    pick = np.sort(rng.choice(ps.B, min(subset, ps.B), replace=False))
    off = ps.dir_off.cpu().numpy()
    idx = ps.dir_idx.cpu().numpy()
    sc = res.corr.scores.cpu().numpy().reshape(-1)
    py = res.corr.peak_y.cpu().numpy().reshape(-1)
    px = res.corr.peak_x.cpu().numpy().reshape(-1)
    best = res.corr.best.cpu().numpy()
    P, C = ps.P, eng.crop
    peak = res.corr.peak.cpu().numpy()
    agree, ties, worst, worst_best, used = 0, 0, 0.0, 0.0, 0
    for b in pick:
        d = idx[off[b]:off[b + 1]]
        if len(d) == 0:
            continue
        dd = torch.from_numpy(d.astype(np.int64)).to(dev)
        rows = raw[dd].unpack() if isinstance(raw, kernels.FragmentPlanes) else raw[dd][..., :eng.n // 2 + 1]
        t = rows.to(torch.complex128) * tinv[dd].double()[:, None, None]
        p = ps.spectra_frag[b].unpack().to(torch.complex128) * ps.inv_norm[b].double()[:, None, None]
        q = t[:, None] * p[None]                                    # [Mi, P, n, m]
        crop = 1024.0 * (Wt @ q @ Et).real.reshape(len(d) * P, C * C)
        v, k = crop.max(dim=1)
        v, k = v.cpu().numpy(), k.cpu().numpy()
        g = slice(off[b] * P, off[b + 1] * P)
        worst = max(worst, float(np.abs(sc[g] - v).max() / np.abs(v).max()))
        bw = int(np.argmax(v))
        worst_best = max(worst_best, abs(float(peak[b]) - float(v[bw])) / abs(float(v[bw])))
        gb = off[b] * P + int(best[b])
        same = int(best[b]) == bw and int(py[gb]) * C + int(px[gb]) == int(k[bw])
        agree += int(same)
        if not same:
            chosen = float(crop[int(best[b]), int(py[gb]) * C + int(px[gb])])
            ties += int(v[bw] - chosen <= 1e-5 * abs(v[bw]))
        used += 1
    return {"subset": int(used), "seed": seed, "top1_agree_fp64": agree / max(used, 1),
            "top1_differs_outside_near_ties": used - agree - ties, "max_rel_err_fp64": worst_best,
            "max_rel_err_fp64_all_pairs": worst,
            "reference": "same operator (every coefficient, the full crop) in float64 (torch complex128)"}


def fp16_variant(eng, ps, tk, res_bf16, subset: int, seed: int, repeats: int = 6):
    """The split-fp16 operand variant of the small-shift search on the same
    template and particle coefficients: search time (HIP events around each
    launch on the launch stream, median of ``repeats`` launches per operand,
    split-bf16 and split-fp16 alternating so clock drift hits both alike),
    scores/s of the search stage, top-1 against the split-bf16 search and, on
    the seeded subset, against the float64 functional (precision_check)."""
    import torch
    times = {"bf16": [], "fp16": []}
    res_ = None
    try:
        for rep in range(repeats + 1):  # the first round is a warm-up
            for operand in ("bf16", "fp16"):
                eng.set_small_shift_operand(operand)
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                res = eng.search_small_shift(None, ps, tk=tk)
                e1.record()
                torch.cuda.synchronize()
                if rep:
                    times[operand].append(e0.elapsed_time(e1))
                if operand == "fp16":
                    res_ = res
        eng.set_small_shift_operand("fp16")
        ms, ms_bf16 = float(np.median(times["fp16"])), float(np.median(times["bf16"]))
        out = {"operand": "fp16 x3 (split hi+lo, A scaled 2^14, fp32 accumulate)", "search_ms": ms,
               "bf16_search_ms": ms_bf16, "search_scores_per_s": ps.n_scores / (ms * 1e-3),
               "top1_same_as_bf16": float(np.mean(res_.corr.best.cpu().numpy() == res_bf16.corr.best.cpu().numpy()))}
        if subset > 0:
            out["precision"] = precision_check(eng, ps, tk, res_, subset, seed)
        return out
    finally:
        eng.set_small_shift_operand("bf16")


def profiled_traffic(kernel: str, workload: str):
    """HBM bytes per launch of ``kernel`` on ``workload`` from the committed
    rocprofv3 PMC summary profiles/<workload>_<kernel>_traffic.md (its
    ``traffic_bytes_per_launch:`` line), or None when there is none."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"{workload}_{kernel}_traffic.md")
    if not os.path.exists(path):
        return None
    with open(path, "r", encoding="utf-8") as fh:
        for line in fh:
            if line.strip().startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1].strip())
    return None


def make_engine(wl, data, dev, resolution):
    from gisapr_amd.engine import SearchEngine
    eng = SearchEngine(box=wl.box, apix=wl.apix, crop=wl.crop, psi_step=wl.psi_step,
                       local_range=wl.local_range, kk=wl.kk, voltage=wl.voltage, cs=wl.cs,
                       mask_radius=wl.mask_radius, mask_edge=wl.mask_edge, fsc_values=data["fsc"],
                       dir_rot=data["dir_rot"], dir_tilt=data["dir_tilt"], device=dev)
    eng.enable_small_shift(resolution=resolution, half=3)
    return eng


def pso_leg(args, dev, log):
    """Batched PSO step (box192 workload): a swarm of ``--pso-slots`` candidates x
    ``--pso-iters`` iterations, poses from a seeded normal distribution (rotation
    capped at 5 deg, translation at 2 A), each swarm member scoring its own
    seeded particle subset of 2,000-20,000 particles; one graph-captured
    batched step per iteration on the ``--pso-path`` search path."""
    import torch
    from gisapr_amd import synthetic
    from gisapr_amd.batch import Autotuner, CandidateBatch, RigidBody
    t0 = time.perf_counter()
    wl = synthetic.WORKLOADS["box192"]
    data = synthetic.make_workload("box192", n_particles=args.pso_particles, device=dev)
    eng = make_engine(wl, data, dev, args.resolution)
    dfo = np.stack([data["dfu"], data["dfv"], data["dfa"]], 1)
    ps = eng.add_particles(data["images"], data["rot"], data["tilt"], data["psi"], defocus=dfo, prepare=False)
    del data["images"]
    if args.pso_path == "mixed":
        eng.prepare_small_shift(ps, drop_spectra=True)
    else:
        eng.prepare(ps)
    rng = np.random.default_rng(synthetic.SEED_CAND + 3)
    S = args.pso_slots
    lo_n, hi_n = min(2000, ps.B), ps.B
    subsets = [np.sort(rng.choice(ps.B, int(rng.integers(lo_n, hi_n + 1)), replace=False)) for _ in range(S)]
    body = RigidBody.from_arrays(data["coords"], data["numbers"], data["chains"], wl.moving_chain)
    batch = CandidateBatch(eng, ps, body, wl.res, S, path=args.pso_path, subsets=subsets, autotuner=Autotuner())

    def swarm():
        rot = rng.normal(0.0, 2.0, (S, 3))
        nr = np.linalg.norm(rot, axis=1, keepdims=True)
        rot = np.where(nr > 5.0, rot * 5.0 / np.maximum(nr, 1e-12), rot)
        tr = rng.normal(0.0, 1.0, (S, 3))
        nt = np.linalg.norm(tr, axis=1, keepdims=True)
        tr = np.where(nt > 2.0, tr * 2.0 / np.maximum(nt, 1e-12), tr)
        return np.concatenate([rot, tr], 1)

    first = swarm()
    batch.tune(first)
    batch.capture(first)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    log("pso setup", setup)
    best = -np.inf
    t1 = time.perf_counter()
    for it in range(args.pso_iters):
        scores = batch.evaluate(swarm())
        best = max(best, float(scores.max()))
    dt = time.perf_counter() - t1
    pairs = batch.plan.n_pairs * args.pso_iters
    return {"workload": f"box192: {sum(wl.chain_sizes)}-atom model (moving chain {wl.chain_sizes[-1]}), "
                        f"{ps.B} particles 192x192 at {wl.apix} A/px, {len(eng.dir_rot)} directions, {ps.P} psi; "
                        f"{S} candidates x {args.pso_iters} iterations, subsets {lo_n}-{hi_n} particles",
            "scores_per_s": pairs / dt, "candidates_per_s": S * args.pso_iters / dt,
            "ms_per_iteration": dt * 1000.0 / args.pso_iters, "pairs_per_iteration": batch.plan.n_pairs,
            "mean_subset": float(np.mean([len(x) for x in subsets])), "tile": dict(batch.tile), "graph": True,
            "path": args.pso_path, "setup_s": setup, "best_score": best}


def serve_leg(args, dev, log):
    """Serving legs: three resident datasets, a seeded Poisson request stream per
    path.  The datasets are built once; the FP32 and full-plane split paths
    share their resident spectra (split16 adds the fp16 copy), and the
    small-shift path (when asked for) runs last because it replaces them by
    their in-disk coefficients.
    Returns {path: stats}."""
    import torch
    from gisapr_amd import synthetic
    from gisapr_amd.batch import Autotuner, RigidBody
    from gisapr_amd.serve import SearchServer, poisson_requests
    t0 = time.perf_counter()
    names = ["serve128", "serve192", "serve256"]
    built = []
    for name in names:
        wl = synthetic.WORKLOADS[name]
        data = synthetic.make_workload(name, device=dev)
        eng = make_engine(wl, data, dev, args.resolution)
        dfo = np.stack([data["dfu"], data["dfv"], data["dfa"]], 1)
        ps = eng.add_particles(data["images"], data["rot"], data["tilt"], data["psi"], defocus=dfo)
        del data["images"]
        body = RigidBody.from_arrays(data["coords"], data["numbers"], data["chains"], wl.moving_chain)
        built.append((name, wl, eng, ps, body))
        log(f"serving dataset {name} built")
    setup = time.perf_counter() - t0
    requests = {"fp32": args.serve_requests_fp32, "split": args.serve_requests, "split16": args.serve_requests,
                "mixed": args.serve_requests}
    out = {}
    for path in ("fp32", "split", "split16", "mixed"):
        if requests[path] <= 0 or path not in args.serve_paths.split(","):
            continue
        t1 = time.perf_counter()
        server = SearchServer(dev, path=path, slots=args.serve_slots, autotuner=Autotuner())
        sizes = {}
        for name, wl, eng, ps, body in built:
            ds = server.add_dataset(name, eng, ps, body, wl.res)
            sizes[name] = {"box": wl.box, "particles": ps.B, "resident_GiB": ds.resident_bytes / 2 ** 30}
            log(f"serving dataset {name} resident on the {path} path ({ds.resident_bytes / 2**30:.1f} GiB)")
        # warm every executor once (library set-up, autotuning) before the stream
        for name in names:
            for ex in server.datasets[name].executors:
                ex.tune(np.zeros((server.slots, 6)))
                ex.evaluate(np.zeros((1, 6)))
        torch.cuda.synchronize()
        reqs = poisson_requests(requests[path], args.serve_rate, names, seed=7)
        stats = server.serve(reqs)
        stats.update(datasets=sizes, setup_s=setup + time.perf_counter() - t1 - stats["wall_s"],
                     rate_per_s=args.serve_rate, slots=server.slots, path=path,
                     offered_candidates_per_s=args.serve_rate * float(np.mean([r.size for r in reqs])))
        out[path] = stats
        log(f"serving {path}: p50 {stats['p50_ms']:.0f} ms, p99 {stats['p99_ms']:.0f} ms")
        del server
        torch.cuda.empty_cache()
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="box256")
    ap.add_argument("--particles", type=int, default=None)
    ap.add_argument("--slots", type=int, default=8, help="candidates per batched step")
    ap.add_argument("--resolution", type=float, default=4.0, help="small-shift disk radius in A")
    ap.add_argument("--precision-subset", type=int, default=256)
    ap.add_argument("--no-fp16-variant", dest="fp16_variant", action="store_false",
                    help="skip the split-fp16 operand variant of the small-shift search")
    ap.add_argument("--cpu-sample", type=int, default=100, help="box256 particles timed on the CPU (0 to skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--pso-iters", type=int, default=20, help="batched PSO leg iterations (0 to skip)")
    ap.add_argument("--pso-slots", type=int, default=64)
    ap.add_argument("--pso-particles", type=int, default=20000)
    ap.add_argument("--serve-requests", type=int, default=2000,
                    help="split / split16 serving stream (0 skips the serving legs)")
    ap.add_argument("--serve-requests-fp32", type=int, default=40, help="FP32 serving stream (0 to skip it)")
    ap.add_argument("--serve-paths", default="fp32,split16")
    ap.add_argument("--pso-path", default="split16", help="search path of the batched PSO leg")
    ap.add_argument("--serve-rate", type=float, default=50.0)
    ap.add_argument("--serve-slots", type=int, default=16)
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="after the legs, run one device-to-device copy of this many MiB (a known-size kernel "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--legs", default="fp32,fft32,split,split16,mixed,batched_fp32,batched_split,batched_split16",
                    help="comma list of legs to run (profiling runs select one)")
    ap.add_argument("--verbose", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: the engine runs on one GPU only; --gpus must be 1", file=sys.stderr)
        return 2

    def log(*a):
        if args.verbose:
            print("[bench]", *a, file=sys.stderr, flush=True)

    import torch
    from gisapr_amd import synthetic
    from gisapr_amd.batch import Autotuner, CandidateBatch, RigidBody
    from gisapr_amd.pdb_text import move_chain

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    wl = synthetic.WORKLOADS[args.workload]
    t_setup = time.perf_counter()
    data = synthetic.make_workload(args.workload, n_particles=args.particles, device=dev)
    eng = make_engine(wl, data, dev, args.resolution)
    tab = eng.small
    dfo = np.stack([data["dfu"], data["dfv"], data["dfa"]], 1)
    ps = eng.add_particles(data["images"], data["rot"], data["tilt"], data["psi"], defocus=dfo, prepare=False)
    del data["images"]
    torch.cuda.synchronize()
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    e0.record()
    eng.prepare(ps)
    e1.record()
    eng.prepare_small_shift(ps)
    e2.record()
    torch.cuda.synchronize()
    prep_ms, gather_ms = e0.elapsed_time(e1), e1.elapsed_time(e2)
    n_scores = ps.n_scores
    rng = np.random.default_rng(synthetic.SEED_CAND)
    npose = max(args.warmup + args.steps, 1) * max(args.slots, 1)
    poses = np.stack([synthetic.candidate_pose(seed=int(s), rot_deg=wl.cand_rot_deg, trans_A=wl.cand_trans_A)
                      for s in rng.integers(0, 2 ** 31, npose)])
    coords, numbers, chains = data["coords"], data["numbers"], data["chains"]
    setup_s = time.perf_counter() - t_setup
    log("setup", setup_s)

    def run_single(make_templates, search):
        stages = ("density", "templates", "search")
        ev_time = {k: 0.0 for k in stages}

        def step(pose, timed):
            xyz = move_chain(coords, chains, wl.moving_chain, pose)
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)] if timed else None
            if timed:
                ev[0].record()
            vol = eng.density(xyz, numbers, wl.res)
            if timed:
                ev[1].record()
            tpl = make_templates(vol)
            if timed:
                ev[2].record()
            res = search(tpl, ps)
            if timed:
                ev[3].record()
            return res, tpl, ev

        for i in range(args.warmup):
            step(poses[i], False)
        torch.cuda.synchronize()
        evs = []
        t0 = time.perf_counter()
        for i in range(args.steps):
            res, tpl, ev = step(poses[args.warmup + i], True)
            evs.append(ev)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        for ev in evs:
            for k, name in enumerate(stages):
                ev_time[name] += ev[k].elapsed_time(ev[k + 1])
        return {"scores_per_s": n_scores * args.steps / dt, "ms_per_candidate": dt * 1000.0 / args.steps,
                "stage_ms": {k: v / args.steps for k, v in ev_time.items()},
                "score": float(res.score.cpu())}, res, tpl

    want = set(args.legs.split(","))
    legs, results = {}, {}
    bank = bank_split = bank_split16 = tk_last = None  # last candidate's (bank, 1/norms) / template coefficients
    if "fp32" in want:
        # FP32 path: full unnormalised bank (1/norm folded into the search)
        legs["fp32"], results["fp32"], bank = run_single(lambda v: eng.templates(v, normalize=False, fragments=True),
                                                         lambda t, p: eng.search(t[0], p, t[1]))
    if "split" in want:
        # the same search on split-bf16 MFMA, the same bank values
        legs["split"], results["split"], bank_split = run_single(
            lambda v: eng.templates(v, normalize=False, fragments=True), lambda t, p: eng.search_split(t[0], p, t[1]))
    if "split16" in want:
        # fp16 operands, then the best pair and near ties of each particle on split-bf16
        legs["split16"], results["split16"], bank_split16 = run_single(
            lambda v: eng.templates(v, normalize=False, fragments=True),
            lambda t, p: eng.search_split(t[0], p, t[1], operand="fp16"))
    if "mixed" in want:
        # split-bf16 path: the template kernel writes the in-disk coefficients directly
        legs["mixed"], results["mixed"], tk_last = run_single(eng.small_templates,
                                                              lambda t, p: eng.search_small_shift(None, p, tk=t))
    if "fft32" in want and "fp32" in legs:
        # the reference's own formulation of the same search (irfft2 per pair) on the same bank, timed alone
        t_bank, t_inv = bank
        fft32_search(eng, ps, t_bank, t_inv)  # warm-up (rocFFT plans, the row-major spectra)
        torch.cuda.synchronize()
        reps = max(1, min(args.steps, 2))
        t0 = time.perf_counter()
        for _ in range(reps):
            fsc_, fy, fx, fbest, fpeak = fft32_search(eng, ps, t_bank, t_inv)
        torch.cuda.synchronize()
        ms_fft = (time.perf_counter() - t0) * 1000.0 / reps
        ref = results["fp32"]
        b32, bf = ref.corr.best.cpu().numpy(), fbest.cpu().numpy()
        off = ps.dir_off.cpu().numpy()
        py32, px32 = ref.corr.peak_y.cpu().numpy().reshape(-1), ref.corr.peak_x.cpu().numpy().reshape(-1)
        fyn, fxn = fy.cpu().numpy(), fx.cpu().numpy()
        used = same_o = same_os = 0
        for i in range(ps.B):
            if b32[i] < 0:
                continue
            used += 1
            if b32[i] == bf[i]:
                same_o += 1
                g = off[i] * ps.P + b32[i]
                same_os += int(py32[g] == fyn[g] and px32[g] == fxn[g])
        m = b32 >= 0
        p32 = ref.corr.peak.cpu().numpy()[m]
        dens_tpl = legs["fp32"]["stage_ms"]["density"] + legs["fp32"]["stage_ms"]["templates"]
        legs["fft32"] = {
            "formulation": "v606 literally: pre-shifted product, torch.fft.irfft2 (rocFFT, float32) per pair, "
                           "C x C crop, max; torch on the GPU, 1,024 pairs per chunk (a baseline, not a product path)",
            "search_ms": ms_fft, "search_scores_per_s": n_scores / (ms_fft * 1e-3),
            "ms_per_candidate": dens_tpl + ms_fft, "scores_per_s": n_scores / ((dens_tpl + ms_fft) * 1e-3),
            "candidate_note": "density + templates from the fp32 leg's stage times, plus this search",
            "top1_same_as_fp32": {"orientation": same_o / max(used, 1), "orientation_and_shift": same_os / max(used, 1),
                                  "particles": used},
            "max_rel_peak_diff_vs_fp32": float(np.max(np.abs(fpeak.cpu().numpy()[m] - p32) / np.abs(p32))),
        }
        del fsc_, fy, fx, fbest, fpeak
        ps.spectra = None  # the row-major copy was for this leg only
        torch.cuda.empty_cache()
    log("single legs", {k: v["ms_per_candidate"] for k, v in legs.items()})

    # batched legs: ``slots`` candidates per step, graph-captured, autotuned search tile
    body = RigidBody.from_arrays(coords, numbers, chains, wl.moving_chain)
    tuner = Autotuner()
    for path in ("fp32", "split", "split16", "mixed"):
        if "batched_" + path not in want:
            continue
        batch = CandidateBatch(eng, ps, body, wl.res, args.slots, path=path, autotuner=tuner)
        batch.tune(poses[:args.slots])
        batch.capture(poses[:args.slots])
        for i in range(args.warmup):
            batch.evaluate(poses[i * args.slots:(i + 1) * args.slots])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(args.steps):
            j = (args.warmup + i) * args.slots
            out = batch.evaluate(poses[j:j + args.slots])
        dt = time.perf_counter() - t0
        pairs = batch.plan.n_pairs
        legs["batched_" + path] = {"scores_per_s": pairs * args.steps / dt,
                                   "ms_per_step": dt * 1000.0 / args.steps,
                                   "ms_per_candidate": dt * 1000.0 / (args.steps * args.slots),
                                   "slots": args.slots, "tile": dict(batch.tile), "graph": True,
                                   "last_scores": [float(x) for x in out[:2]]}
        del batch
        torch.cuda.empty_cache()
    log("batched legs", {k: v["ms_per_candidate"] for k, v in legs.items()})

    # roofline of each single leg's search stage (HIP events on the launch stream)
    n = eng.n
    if "fp32" in legs:
        ms = legs["fp32"]["stage_ms"]["search"]
        ach = correlation_flops_per_score(n) * n_scores / (ms * 1e-3) / 1e12
        legs["fp32"]["roofline"] = {"bound": "mfma",
                                    "kernel": "split_correlate_kernel, FP32 operands (gis_correlate; + tables, tail)",
                                    "achieved": ach, "peak": PEAK_FP32_MFMA_TFLOPS, "unit": "TFLOP/s",
                                    "frac": ach / PEAK_FP32_MFMA_TFLOPS,
                                    "flops_per_score": correlation_flops_per_score(n), "launch_ms": ms,
                                    "traffic": profiled_traffic("fp32_search", wl.name)}
    if "split" in legs:
        ms = legs["split"]["stage_ms"]["search"]
        fps = split_flops_per_score(n, eng.crop)
        ach = fps * n_scores / (ms * 1e-3) / 1e12
        legs["split"]["roofline"] = {"bound": "mfma", "kernel": "split_correlate_kernel (+ tables, particle tail)",
                                     "achieved": ach, "peak": PEAK_BF16_MFMA_TFLOPS, "unit": "TFLOP/s",
                                     "frac": ach / PEAK_BF16_MFMA_TFLOPS, "flops_per_score": fps,
                                     "launch_ms": ms,
                                     "traffic": profiled_traffic("split_search", wl.name)}
    for leg in ("split", "split16"):
        if leg not in legs:
            continue
        ms_t = legs[leg]["stage_ms"]["templates"]
        bpd = template_bytes_per_direction(eng.N, n)
        ach_t = bpd * len(eng.dir_rot) / (ms_t * 1e-3) / 1e9
        legs[leg]["roofline_templates"] = {
            "bound": "hbm", "kernel": "volume transform + build_templates_kernel", "achieved": ach_t,
            "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": ach_t / PEAK_HBM_GBS,
            "bytes_per_direction": bpd, "directions": len(eng.dir_rot), "launch_ms": ms_t,
            "traffic": profiled_traffic("templates_stage", wl.name)}
    if "split16" in legs:
        ms = legs["split16"]["stage_ms"]["search"]
        fps = split16_flops_per_score(n, eng.crop)
        ach = fps * n_scores / (ms * 1e-3) / 1e12
        legs["split16"]["roofline"] = {
            "bound": "mfma", "kernel": "split16_correlate_kernel (+ tables, particle tail, split-bf16 rescoring)",
            "achieved": ach, "peak": PEAK_BF16_MFMA_TFLOPS, "unit": "TFLOP/s",
            "frac": ach / PEAK_BF16_MFMA_TFLOPS, "flops_per_score": fps, "launch_ms": ms,
            "traffic": profiled_traffic("split16_search", wl.name)}
    if "mixed" in legs:
        ms = legs["mixed"]["stage_ms"]["search"]
        fps = mixed_flops_per_score(tab.K, tab.n_lags)
        ach = fps * n_scores / (ms * 1e-3) / 1e12
        legs["mixed"]["roofline"] = {"bound": "mfma",
                                     "kernel": "mixed_correlate_kernel (+ particle tail)",
                                     "achieved": ach, "peak": PEAK_BF16_MFMA_TFLOPS, "unit": "TFLOP/s",
                                     "frac": ach / PEAK_BF16_MFMA_TFLOPS, "flops_per_score": fps,
                                     "launch_ms": ms, "traffic": profiled_traffic("mixed_search", wl.name)}

    if args.copy_calibration_mb > 0:
        src = torch.ones(args.copy_calibration_mb << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    head_name = next(k for k in ("batched_split16", "batched_split", "split16", "split", "batched_mixed", "mixed",
                                 "batched_fp32", "fp32") if k in legs)
    head = legs[head_name]
    single = next((k for k in ("split16", "split", "mixed", "fp32") if k in legs), None)
    roof = dict(legs[single].get("roofline", {})) if single else None  # no single-candidate leg: no roofline
    if single in ("split", "split16") and legs[single].get("roofline_templates"):
        # the line's roofline is the kernel that takes the larger share of a candidate step
        st = legs[single]["stage_ms"]
        if st["templates"] > st["search"]:
            roof = dict(legs[single]["roofline_templates"])
        roof["from_leg"] = f"{single} (one candidate per step, HIP events on the launch stream)"
    out = {
        "metric": METRIC, "value": head["scores_per_s"], "unit": "scores/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head.get("ms_per_step", head["ms_per_candidate"]),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": ("fp16 operands (packed fp16 products, fp32 accumulate) + bf16x3 rescoring of each particle's "
                  "best pair and near ties" if "split16" in head_name
                  else "bf16x3 (split hi+lo, fp32 accumulate)" if ("mixed" in head_name or "split" in head_name)
                  else "fp32"),
        "data": "synthetic (seeded generators; random-walk model)",
        "config": {"workload": f"{wl.name}: {sum(wl.chain_sizes)}-atom model (moving chain "
                               f"{wl.chain_sizes[-1]}), {ps.B} particles {wl.box}x{wl.box} at {wl.apix} A/px, "
                               f"HEALPix order {wl.healpix_order} ({len(eng.dir_rot)} directions) local "
                               f"+/-{wl.local_range} deg, psi step {wl.psi_step} ({ps.P} psi); {args.slots} "
                               f"candidates per batched step, particle spectra resident",
                   "box": wl.box, "particles": ps.B, "directions": len(eng.dir_rot), "psi": ps.P,
                   "scores_per_candidate": n_scores, "path": head_name, "slots": args.slots,
                   "crop": wl.crop, "mixed_lags": tab.n_lags, "mixed_K": tab.K,
                   "mixed_resolution_A": args.resolution},
        "candidates_per_s": 1000.0 / head["ms_per_candidate"],
        "particle_prep_ms": prep_ms, "coefficient_gather_ms": gather_ms,
        "legs": legs,
        "setup_s": setup_s,
        "roofline": roof,
        "cpu_baseline": None,
    }
    if "fp32" in legs:
        out["speedup_vs_fp32"] = {k: v["scores_per_s"] / legs["fp32"]["scores_per_s"] for k, v in legs.items()}
    if "fft32" in legs:
        f = legs["fft32"]
        out["speedup_vs_fft32"] = {
            "candidate": {k: v["scores_per_s"] / f["scores_per_s"] for k, v in legs.items()},
            "search": {k: f["search_ms"] / v["stage_ms"]["search"] for k, v in legs.items() if "stage_ms" in v},
            "note": "against the reference's own formulation of the FP32 search (irfft2 per pair, torch/rocFFT)"}
    agreement = {}
    for other, center in (("split", eng.crop // 2), ("split16", eng.crop // 2), ("mixed", 3)):
        if "fp32" in results and other in results:
            a = top1(results["fp32"], ps, eng.crop // 2)
            b = top1(results[other], ps, center)
            valid = a[:, 0] >= 0
            agreement[other] = {
                "orientation": float(np.mean((a[valid, 0] == b[valid, 0]))),
                "orientation_and_shift": float(np.mean(np.all(a[valid] == b[valid], axis=1))),
                "particles": int(valid.sum())}
    if "split16" in agreement:
        out["top1_same_as_fp32"] = dict(agreement["split16"], path="split16 (same operator and crop)")
    if "split" in agreement:
        out["top1_split_bf16_vs_fp32"] = dict(agreement["split"], path="split (same operator and crop)")
    if "split16" in results:
        # the split-bf16 search on the split16 leg's own last bank: the rescored
        # answers are that search's, bit for bit
        a = eng.search_split(bank_split16[0], ps, bank_split16[1]).corr
        b = results["split16"].corr
        valid = a.best >= 0
        g = (ps.dir_off[:-1].long() * ps.P + a.best.long().clamp(min=0))
        same_lag = (a.peak_y.reshape(-1)[g] == b.peak_y.reshape(-1)[g]) & \
                   (a.peak_x.reshape(-1)[g] == b.peak_x.reshape(-1)[g])
        out["split16_vs_split"] = {
            "same_best": float((a.best == b.best)[valid].float().mean()),
            "same_best_lag": float(same_lag[valid].float().mean()),
            "bit_identical_peak": float((a.peak == b.peak)[valid].float().mean()),
            "particles": int(valid.sum()), "bank": "split16 leg's last candidate"}
    if "mixed" in agreement:
        out["top1_small_shift_vs_fp32"] = dict(
            agreement["mixed"], path="mixed small-shift variant (truncated coefficients, +/-3 px lags)")
    if args.precision_subset > 0 and "split16" in results:
        out["precision"] = precision_check_split(eng, ps, bank_split16, results["split16"], args.precision_subset,
                                                 synthetic.SEED_CAND)
    if args.precision_subset > 0 and "split" in results:
        out["precision_split_bf16"] = precision_check_split(eng, ps, bank_split, results["split"],
                                                            args.precision_subset, synthetic.SEED_CAND)
    if args.precision_subset > 0 and "mixed" in results:
        out["precision_small_shift"] = precision_check(eng, ps, tk_last, results["mixed"], args.precision_subset,
                                                       synthetic.SEED_CAND)
    if args.fp16_variant and "mixed" in results:
        out["fp16_variant"] = fp16_variant(eng, ps, tk_last, results["mixed"], args.precision_subset,
                                           synthetic.SEED_CAND)
    if not args.no_cpu_baseline and args.cpu_sample > 0:
        out["cpu_baseline"] = cpu_baseline(n_scores, args.cpu_sample, log=log)
    del ps, eng, bank, bank_split, bank_split16, tk_last, results
    torch.cuda.empty_cache()
    if args.pso_iters > 0:
        out["pso"] = pso_leg(args, dev, log)
        torch.cuda.empty_cache()
    if args.serve_requests > 0:
        served = serve_leg(args, dev, log)
        main_path = next((p for p in ("split16", "split", "mixed") if p in served), None)
        if main_path is not None:
            out["serving"] = served[main_path]
            out["p50_ms"] = served[main_path]["p50_ms"]
            out["p99_ms"] = served[main_path]["p99_ms"]
        if "fp32" in served:
            out["serving_fp32"] = served["fp32"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
