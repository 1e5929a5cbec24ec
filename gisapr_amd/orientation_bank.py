"""Per-particle Top-K direction banks for calibrated searches.

A calibration search ranks every particle's local directions by their best
score over psi and stores the first ``keep`` directions with distinct
(rot, tilt); later searches score only those.  Used by run_search's
``orientation_rows`` / ``calibrate`` arguments and ``--orientations-in`` /
``--orientations-out`` flags (the zip release's calibration mode).

On disk the bank is one JSON object (this layout and the digest rule are a
file format shared with the zip release, so banks move between the two):

    {"settings": {"schema", "particle_sha256", "angle_sha256", "keep", "local",
                  "local_range", "psi_step"},
     "rows": {"<STAR line>": {"image": "<image name>",
                              "directions": [[index, rot, tilt, score], ...]}},
     ...origin fields..., "sha256": "<digest>"}

The digest is SHA-256 of the object without "sha256", serialised as JSON
with sorted keys and no whitespace.  A bank is refused when its digest or
its settings do not match the current inputs.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import math
import os
import tempfile
from functools import lru_cache
from pathlib import Path
from typing import Dict, List, Sequence

import numpy as np

SCHEMA = "gisapr-orientations-v1"

__all__ = ["positive_int", "file_key", "file_hash", "settings", "pack_bank", "bank_digest", "check_bank",
           "restricted_indices", "ranked_row", "atomic_json", "SCHEMA"]


def positive_int(text) -> int:
    """argparse type of --keep-orientations: an integer of at least 1."""
    value = int(text)
    if value >= 1:
        return value
    raise argparse.ArgumentTypeError(f"{text!r}: must be >= 1")


def file_key(path) -> tuple:
    """What identifies a file's current contents without reading it: resolved
    path, size, modification and change times (ns)."""
    resolved = Path(path).expanduser().resolve()
    info = resolved.stat()
    return str(resolved), info.st_size, info.st_mtime_ns, info.st_ctime_ns


@lru_cache(maxsize=16)
def _digest_of_file(key: tuple) -> str:
    h = hashlib.sha256()
    with open(key[0], "rb") as fh:
        while True:
            chunk = fh.read(1 << 20)
            if not chunk:
                return h.hexdigest()
            h.update(chunk)


def file_hash(path) -> str:
    """SHA-256 of a file's bytes, computed once per file_key."""
    return _digest_of_file(file_key(path))


def settings(particles, angles, keep: int, local: bool, angular_range: float, psi_step: float) -> Dict:
    """The inputs and search settings a bank belongs to."""
    return dict(schema=SCHEMA, particle_sha256=file_hash(particles), angle_sha256=file_hash(angles),
                keep=int(keep), local=bool(local), local_range=float(angular_range), psi_step=float(psi_step))


def bank_digest(bank: Dict) -> str:
    """The bank's digest: SHA-256 of its canonical JSON without "sha256"."""
    canonical = json.dumps({k: bank[k] for k in bank if k != "sha256"}, sort_keys=True, separators=(",", ":"),
                           allow_nan=False)
    return hashlib.sha256(canonical.encode("utf-8")).hexdigest()


def pack_bank(context: Dict, rows: Dict, **origin) -> Dict:
    """A bank of ``rows`` for ``context`` (settings), with its digest."""
    bank = dict(origin, settings=context, rows=rows)
    bank["sha256"] = bank_digest(bank)
    return bank


def check_bank(bank: Dict, context: Dict) -> None:
    """ValueError unless the bank was built for ``context`` and is unaltered."""
    if bank.get("settings") == context and bank.get("sha256") == bank_digest(bank):
        return
    raise ValueError("orientation bank does not match the inputs, K or search settings (or was altered)")


def restricted_indices(row, rec, allowed: Sequence[int], angles, keep: int) -> List[int]:
    """The direction indices a bank row lets particle ``rec`` search, in
    ascending order (ties then fall in angle-file order).

    The row must name the particle's image, list at most min(keep, number of
    local directions) entries -- none only when the particle has no local
    direction -- and each entry must be a local direction index whose
    (rot, tilt) equals the angle table's, appear once by (rot, tilt), and
    carry a finite score."""
    where = f"STAR line {rec.line_index}"
    if not isinstance(row, dict) or row.get("image") != rec.image_name:
        raise ValueError(f"orientation bank has no entry for the image at {where}")
    local = {int(i) for i in allowed}
    entries = row["directions"]
    if len(entries) > min(int(keep), len(local)) or (local and not entries):
        raise ValueError(f"orientation bank lists {len(entries)} directions at {where}")
    picked: Dict[tuple, int] = {}
    for index, rot, tilt, score in entries:
        valid = (type(index) is int and index in local and (rot, tilt) == (angles.rot[index], angles.tilt[index])
                 and (rot, tilt) not in picked and math.isfinite(score))
        if not valid:
            raise ValueError(f"orientation bank entry {[index, rot, tilt, score]} is not valid at {where}")
        picked[(rot, tilt)] = index
    return sorted(picked.values())


def ranked_row(rec, indices: Sequence[int], angles, scores, keep: int) -> Dict:
    """Bank row of one particle from its score table ``scores`` [len(indices),
    P] (each entry already the best over translations): directions ranked by
    their best score over psi, descending, ties in table order, keeping the
    first ``keep`` with distinct (rot, tilt).  A NaN best score marks the row
    invalid (no directions); an infinite one is an error."""
    if len(indices) == 0:
        return {"image": rec.image_name, "directions": []}
    table = scores.detach().cpu().numpy() if hasattr(scores, "detach") else scores
    best = np.asarray(table, np.float64).reshape(len(indices), -1).max(axis=1)
    if np.isinf(best).any():
        raise ValueError(f"infinite directional peak at STAR line {rec.line_index}")
    if np.isnan(best).any():
        return {"image": rec.image_name, "directions": [], "invalid_rawcc": True}
    kept: Dict[tuple, list] = {}
    for j in np.argsort(-best, kind="stable"):
        if len(kept) == int(keep):
            break
        i = int(indices[j])
        key = (angles.rot[i], angles.tilt[i])
        kept.setdefault(key, [i, key[0], key[1], float(best[j])])
    return {"image": rec.image_name, "directions": list(kept.values())}


def atomic_json(path, obj) -> None:
    """Write ``obj`` as JSON so readers never see a partial file: into a
    temporary file in the same folder, then renamed over ``path``."""
    target = Path(path)
    target.parent.mkdir(parents=True, exist_ok=True)
    fd, tmp = tempfile.mkstemp(dir=target.parent, prefix=target.name + ".", suffix=".tmp")
    try:
        with os.fdopen(fd, "w", encoding="utf-8") as fh:
            json.dump(obj, fh, allow_nan=False, ensure_ascii=False, indent=2)
            fh.write("\n")
        os.replace(tmp, target)
    except BaseException:
        if os.path.exists(tmp):
            os.unlink(tmp)
        raise
