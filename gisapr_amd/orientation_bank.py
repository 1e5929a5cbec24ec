"""Per-particle Top-K direction banks (the zip release's calibration mode,
ZL orientation_bank_v1.py; used by run_search's ``orientation_rows`` /
``calibrate`` and ``--orientations-in`` / ``--orientations-out``).

A calibration search ranks, for every particle, its local directions by the
best score over psi and keeps the first ``keep`` distinct (rot, tilt) pairs;
later searches score only those directions.  The bank is JSON:

    {"settings": {...inputs and search settings...},
     "rows": {"<STAR line>": {"image": "...", "directions": [[i, rot, tilt, score], ...]}},
     "mrc": "<model path>", "sha256": "<digest of everything but the digest>"}

and is refused when its digest or settings do not match the current inputs.
The format (field names, digest rule) follows ZL orientation_bank_v1.py so
banks written by either implementation read in the other.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import math
import os
from functools import lru_cache
from pathlib import Path
from tempfile import NamedTemporaryFile
from typing import Dict, List, Sequence

import numpy as np

SCHEMA = "gisapr-orientations-v1"

__all__ = ["positive_int", "file_key", "file_hash", "settings", "pack_bank", "bank_digest", "check_bank",
           "restricted_indices", "ranked_row", "atomic_json", "SCHEMA"]


def positive_int(value) -> int:
    """argparse type: an integer >= 1 (ZL orientation_bank_v1.py:14-18)."""
    number = int(value)
    if number < 1:
        raise argparse.ArgumentTypeError("must be >= 1")
    return number


def file_key(path) -> tuple:
    """Identity of a file's current contents: resolved path, size, mtime and
    ctime in ns (ZL gisapr_cache.py:9-12)."""
    p = Path(path).expanduser().resolve()
    st = p.stat()
    return str(p), st.st_size, st.st_mtime_ns, st.st_ctime_ns


@lru_cache(maxsize=16)
def _sha256_of(identity: tuple) -> str:
    h = hashlib.sha256()
    with open(identity[0], "rb") as fh:
        for block in iter(lambda: fh.read(1 << 20), b""):
            h.update(block)
    return h.hexdigest()


def file_hash(path) -> str:
    """SHA-256 of a file, cached per identity (ZL orientation_bank_v1.py:39-45)."""
    return _sha256_of(file_key(path))


def settings(particles, angles, keep: int, local: bool, angular_range: float, psi_step: float) -> Dict:
    """The inputs and search settings a bank is valid for (:48-51)."""
    return {"schema": SCHEMA, "particle_sha256": file_hash(particles), "angle_sha256": file_hash(angles),
            "keep": int(keep), "local": bool(local), "local_range": float(angular_range),
            "psi_step": float(psi_step)}


def bank_digest(bank: Dict) -> str:
    """SHA-256 of the canonical JSON of every field but the digest (:60-63)."""
    body = {k: v for k, v in bank.items() if k != "sha256"}
    text = json.dumps(body, sort_keys=True, allow_nan=False, separators=(",", ":"))
    return hashlib.sha256(text.encode()).hexdigest()


def pack_bank(context: Dict, rows: Dict, **origin) -> Dict:
    """Bank object with its digest (:54-57)."""
    bank = {"settings": context, "rows": rows, **origin}
    bank["sha256"] = bank_digest(bank)
    return bank


def check_bank(bank: Dict, context: Dict) -> None:
    """Refuse a bank built for other inputs or settings, or altered (:66-68)."""
    if bank.get("settings") != context or bank.get("sha256") != bank_digest(bank):
        raise ValueError("Orientation bank digest or input/K/search settings mismatch")


def restricted_indices(row, rec, allowed: Sequence[int], angles, keep: int) -> List[int]:
    """Direction indices a bank row allows for one particle, validated against
    the particle's STAR image name, its local directions ``allowed``, the
    angle table and ``keep``; ascending, so tie order follows the angle file
    (:71-90)."""
    if row is None or row.get("image") != rec.image_name:
        raise ValueError(f"Missing or mismatched bank entry at STAR line {rec.line_index}")
    directions = row["directions"]
    allowed_set = set(int(i) for i in allowed)
    if (not directions and allowed_set) or len(directions) > min(int(keep), len(allowed_set)):
        raise ValueError(f"Invalid direction count at STAR line {rec.line_index}")
    seen, out = set(), []
    for entry in directions:
        i, rot, tilt, score = entry
        if (type(i) is not int or i not in allowed_set or (rot, tilt) in seen
                or (rot, tilt) != (angles.rot[i], angles.tilt[i]) or not math.isfinite(score)):
            raise ValueError(f"Invalid bank direction at STAR line {rec.line_index}")
        seen.add((rot, tilt))
        out.append(i)
    return sorted(out)


def ranked_row(rec, indices: Sequence[int], angles, scores, keep: int) -> Dict:
    """A particle's bank row from its score table ``scores`` [len(indices), P]
    (translation peak already taken): the best score over psi per direction,
    ranked descending with ties in table order, first ``keep`` distinct
    (rot, tilt) pairs (:93-112).  NaN scores mark the row invalid; an infinite
    score is an error."""
    s = np.asarray(scores.detach().cpu().numpy() if hasattr(scores, "detach") else scores, np.float64)
    peaks = s.reshape(len(indices), -1).max(axis=1) if len(indices) else np.zeros(0)
    if np.isinf(peaks).any():
        raise ValueError(f"Infinite directional peak at STAR line {rec.line_index}")
    if np.isnan(peaks).any():
        return {"image": rec.image_name, "directions": [], "invalid_rawcc": True}
    directions, seen = [], set()
    for j in np.argsort(-peaks, kind="stable"):
        i = int(indices[j])
        pair = (angles.rot[i], angles.tilt[i])
        if pair in seen:
            continue
        seen.add(pair)
        directions.append([i, *pair, float(peaks[j])])
        if len(directions) == int(keep):
            break
    return {"image": rec.image_name, "directions": directions}


def atomic_json(path, obj) -> None:
    """Write JSON through a temporary file in the target folder and rename it
    into place (:21-36): readers never see a partial bank."""
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    with NamedTemporaryFile("w", encoding="utf-8", dir=path.parent, delete=False) as fh:
        tmp = Path(fh.name)
        try:
            json.dump(obj, fh, allow_nan=False, ensure_ascii=False, indent=2)
            fh.write("\n")
        except BaseException:
            fh.close()
            tmp.unlink(missing_ok=True)
            raise
    try:
        os.replace(tmp, path)
    finally:
        tmp.unlink(missing_ok=True)
