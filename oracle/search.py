"""Oracle: local orientation map, per-particle records, the whole search, and
the cross-particle score.

Test infrastructure only (see ``oracle/__init__.py``).
"""
from __future__ import annotations

import math

import numpy as np

from .fourier import centered_full_to_packed, correlate
from .particles import PackedCTF, prepare_particle, psi_sampling_for_record
from .project import (PI_REF, build_templates, euler_angles_to_matrix, expand_radial_values,
                      generate_soft_mask)


def calculate_angular_distance(rot1, tilt1, psi1, rot2, tilt2, psi2):
    """v606:331-348: mean of the three row-vector angles (degrees)."""
    e1 = euler_angles_to_matrix(rot1, tilt1, psi1)
    e2 = euler_angles_to_matrix(rot2, tilt2, psi2)
    total = 0.0
    for i in range(3):
        d = sum(float(a) * float(b) for a, b in zip(e1[i], e2[i]))
        total += math.acos(min(max(d, -1.0), 1.0)) * 180.0 / PI_REF
    return total / 3.0


def local_orientations(prior_rot, prior_tilt, model_rot, model_tilt, local_range):
    """v606:351-378 for one particle: ascending model indices."""
    mr = np.asarray(model_rot)
    mt = np.asarray(model_tilt)
    rd = np.fabs((mr - prior_rot + 180.0) % 360.0 - 180.0)
    td = np.fabs((mt - prior_tilt + 180.0) % 360.0 - 180.0)
    out = []
    for i in np.where((rd <= local_range) & (td <= local_range))[0]:
        d = calculate_angular_distance(float(mr[i]), float(mt[i]), 0.0, prior_rot, prior_tilt, 0.0)
        if abs(d) <= local_range:
            out.append(int(i))
    return out


def make_record(line_index, prior, fake_psi, model_rot, model_tilt, model_indices,
                scores, ys, xs, crop, trans_scale, stats_dtype=np.float64):
    """v606:1452-1516.  ``prior`` = (rot, tilt, psi).  Mean/std in ``stats_dtype``."""
    flat = np.asarray(scores).reshape(-1)
    k = int(np.argmax(flat))
    P = scores.shape[1]
    li, pi = k // P, k % P
    peak = scores[li, pi]
    py, px = int(ys[li, pi]), int(xs[li, pi])
    gi = int(model_indices[li])
    ad = calculate_angular_distance(model_rot[gi], model_tilt[gi], 0.0, prior[0], prior[1], 0.0)
    f = flat.astype(stats_dtype)
    mean = f.mean()
    sigma = f.std()
    z = (stats_dtype(peak) - mean) / sigma
    return {
        "line": line_index, "z": float(z), "l3": 1 if z > 3.0 else 0,
        "accu": 1 if ad < 11.0 else 0, "rot": model_rot[gi], "tilt": model_tilt[gi],
        "angdist": ad,
        "xshift": round(float(crop // 2 - px) / trans_scale),
        "yshift": round(float(crop // 2 - py) / trans_scale),
        "psi_fake": float(fake_psi[pi]), "psi_prior": float((prior[2] + 360.0) % 360.0),
        "rawcc": float(peak), "best_local": li, "best_psi": pi, "peak_y": py, "peak_x": px,
        "mean": float(mean), "std": float(sigma),
    }


def score_sum(rawcc) -> float:
    """new_method_to_fit_the_2nd_Gaussian_PEAK_v4.py:133-134: sum(exp(rawCC/1024)), float64."""
    return float(np.sum(np.exp(np.asarray(rawcc, np.float64) / 1024.0)))


def run_search(volume, model_rot, model_tilt, particles, *, apix, new_box, psi_step,
               local_range, crop, kk, voltage=300.0, cs=2.7, mask_radius=110, mask_edge=6,
               fsc_values=None, discard_mask=False, do_local_search=True, dtype=np.float32,
               return_scores=False):
    """run_search (v606:1594-1825) on in-memory arrays.

    ``particles``: dict with arrays ``images`` [B,N,N], ``rot``, ``tilt``, ``psi``,
    ``dfu``, ``dfv``, ``dfa`` (``dfu`` None -> no CTF columns, CTF = 1).
    Returns a list of record dicts (one per particle with >= 1 local direction).
    """
    N = volume.shape[0]
    n = int(new_box)
    mask = None if discard_mask else generate_soft_mask((N, N), int(mask_radius), int(mask_edge))
    fsc = np.ones(N, np.float32) if fsc_values is None else np.asarray(fsc_values, np.float32)
    packed_fsc = centered_full_to_packed(expand_radial_values(fsc, n))
    bank = build_templates(volume, model_rot, model_tilt, n, mask, packed_fsc, dtype)
    new_apix = N / n * float(apix)
    trans_scale = float(apix) / new_apix
    ctf = PackedCTF(n, new_apix, cs, voltage, kk)
    recs, all_scores = [], []
    B = len(particles["images"])
    for i in range(B):
        pr = (float(particles["rot"][i]), float(particles["tilt"][i]), float(particles["psi"][i]))
        if do_local_search:
            idx = local_orientations(pr[0], pr[1], model_rot, model_tilt, local_range)
        else:
            idx = list(range(len(model_rot)))
        if not idx:
            continue
        fake, eff = psi_sampling_for_record(pr[2], psi_step, do_local_search, local_range)
        if particles.get("dfu") is None:
            c = np.ones((n, n // 2 + 1), np.float32)
        else:
            c = ctf.one(particles["dfu"][i], particles["dfv"][i], particles["dfa"][i])
        pf = prepare_particle(particles["images"][i], eff, n, mask, c, dtype)
        sc, ys, xs = correlate(bank[idx], pf, crop, dtype)
        rec = make_record(i, pr, fake, model_rot, model_tilt, idx, sc, ys, xs, crop, trans_scale)
        recs.append(rec)
        all_scores.append(sc)
    return (recs, all_scores) if return_scores else recs
