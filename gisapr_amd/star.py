"""RELION STAR loops: direction tables and particle tables.

Follows the v606 parsing rules (v606:75-316): the first ``loop_`` whose
labels include the required ones; label columns from ``#n`` (1-based) or
their order; data rows until the next ``loop_`` / ``data_`` / label line;
blank and ``#`` lines skipped.  Rows with too few columns are ignored.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Sequence

import numpy as np


@dataclass
class StarLoop:
    lines: List[str]
    labels: Dict[str, int]
    data_line_indices: List[int]
    data_start: int

    def tokens(self, line_index: int) -> List[str]:
        return self.lines[line_index].split()


def _label_column(tokens: Sequence[str], fallback: int) -> int:
    for tok in tokens[1:]:
        if tok.startswith("#"):
            try:
                return int(tok[1:]) - 1
            except ValueError:
                pass
    return fallback


def find_star_loop(lines: List[str], required_labels: Sequence[str]) -> StarLoop:
    required = set(required_labels)
    n = len(lines)
    i = 0
    while i < n:
        if lines[i].strip().lower() != "loop_":
            i += 1
            continue
        labels: Dict[str, int] = {}
        j, fallback = i + 1, 0
        while j < n:
            s = lines[j].strip()
            if not s or s.startswith("#"):
                j += 1
                continue
            if not s.startswith("_"):
                break
            tok = s.split()
            labels[tok[0]] = _label_column(tok, fallback)
            fallback += 1
            j += 1
        if not required.issubset(labels):
            i = max(j, i + 1)
            continue
        rows: List[int] = []
        k, start = j, -1
        while k < n:
            s = lines[k].strip()
            low = s.lower()
            if low == "loop_" or low.startswith("data_") or s.startswith("_"):
                if rows:
                    break
                k += 1
                continue
            if not s or s.startswith("#"):
                k += 1
                continue
            if start < 0:
                start = k
            rows.append(k)
            k += 1
        if start < 0:
            raise ValueError("Found a STAR loop with the required labels but no data rows: "
                             + ", ".join(required_labels))
        return StarLoop(lines, labels, rows, start)
    raise ValueError("Cannot find a STAR loop containing: " + ", ".join(required_labels))


def read_text_lines(filename: str) -> List[str]:
    with open(filename, "r", encoding="utf-8", errors="replace") as fh:
        return fh.readlines()


@dataclass
class ModelAngles:
    rot: List[float]
    tilt: List[float]
    psi: List[float]
    rotation_matrices_t: np.ndarray  # [M, 3, 3] float32, Euler(rot, tilt, 0)^T


def read_model_angles(angle_star: str) -> ModelAngles:
    """Direction table (v606:190-225); projections ignore the STAR psi."""
    from .host import rotation_matrices_t
    lines = read_text_lines(angle_star)
    loop = find_star_loop(lines, ["_rlnAngleRot", "_rlnAngleTilt"])
    ri, ti = loop.labels["_rlnAngleRot"], loop.labels["_rlnAngleTilt"]
    pi = loop.labels.get("_rlnAnglePsi", -1)
    rot, tilt, psi = [], [], []
    for li in loop.data_line_indices:
        tok = loop.tokens(li)
        if max(ri, ti) >= len(tok):
            continue
        rot.append(float(tok[ri]))
        tilt.append(float(tok[ti]))
        psi.append(float(tok[pi]) if 0 <= pi < len(tok) else 0.0)
    if not rot:
        raise ValueError(f"No model orientations found in {angle_star}")
    return ModelAngles(rot, tilt, psi, rotation_matrices_t(rot, tilt))


@dataclass
class ParticleRecord:
    line_index: int
    image_name: str
    rot: float
    tilt: float
    psi: float
    defocus_u: float = 0.0
    defocus_v: float = 0.0
    defocus_angle: float = 0.0


@dataclass
class ParticleTable:
    lines: List[str]
    loop: StarLoop
    records: List[ParticleRecord]
    records_by_line: Dict[int, ParticleRecord]
    has_ctf: bool


def read_particle_table(particle_star: str) -> ParticleTable:
    """Particle rows with priors and (optionally) defocus (v606:249-297)."""
    lines = read_text_lines(particle_star)
    loop = find_star_loop(lines, ["_rlnImageName", "_rlnAngleRot", "_rlnAngleTilt"])
    lab = loop.labels
    ii, ri, ti = lab["_rlnImageName"], lab["_rlnAngleRot"], lab["_rlnAngleTilt"]
    pi = lab.get("_rlnAnglePsi", -1)
    du, dv, da = lab.get("_rlnDefocusU", -1), lab.get("_rlnDefocusV", -1), lab.get("_rlnDefocusAngle", -1)
    has_ctf = min(du, dv, da) >= 0
    need = max(ii, ri, ti)
    recs: List[ParticleRecord] = []
    for li in loop.data_line_indices:
        tok = loop.tokens(li)
        if need >= len(tok):
            continue
        r = ParticleRecord(li, tok[ii], float(tok[ri]), float(tok[ti]),
                           float(tok[pi]) if 0 <= pi < len(tok) else 0.0)
        if has_ctf:
            r.defocus_u, r.defocus_v, r.defocus_angle = float(tok[du]), float(tok[dv]), float(tok[da])
        recs.append(r)
    if not recs:
        raise ValueError(f"No particle rows found in {particle_star}")
    return ParticleTable(lines, loop, recs, {r.line_index: r for r in recs}, has_ctf)


def select_particle_records(table: ParticleTable, start_line: int, end_line: int) -> List[ParticleRecord]:
    """Rows with start <= line index < end (v606:300-309; -1 = open)."""
    start = max(start_line, table.loop.data_start) if start_line >= 0 else table.loop.data_start
    end = len(table.lines) if end_line < 0 else min(end_line, len(table.lines))
    if end < table.loop.data_start:
        end = table.loop.data_start
    return [r for r in table.records if start <= r.line_index < end]


def write_direction_star(path, rot, tilt, psi=None):
    """A direction table in the layout of generate_healpix_order_and_relion_star.py."""
    psi = np.zeros(len(rot)) if psi is None else psi
    with open(path, "w") as fh:
        fh.write("# relion 30001\n\ndata_particles\n\nloop_\n_rlnAngleRot #1\n_rlnAngleTilt #2\n"
                 "_rlnAnglePsi #3\n_rlnImageName #4\n")
        for r, t, p in zip(rot, tilt, psi):
            fh.write(f"{r}\t{t}\t{p}\t1@1.mrcs\n")


def write_particle_star(path, stack_name, rot, tilt, psi, dfu=None, dfv=None, dfa=None):
    """A particle table referencing images 1..B of one MRC stack."""
    cols = ["_rlnImageName", "_rlnAngleRot", "_rlnAngleTilt", "_rlnAnglePsi"]
    if dfu is not None:
        cols += ["_rlnDefocusU", "_rlnDefocusV", "_rlnDefocusAngle"]
    with open(path, "w") as fh:
        fh.write("# relion 30001\n\ndata_particles\n\nloop_\n")
        for i, c in enumerate(cols):
            fh.write(f"{c} #{i + 1}\n")
        for b in range(len(rot)):
            row = [f"{b + 1}@{stack_name}", repr(float(rot[b])), repr(float(tilt[b])), repr(float(psi[b]))]
            if dfu is not None:
                row += [repr(float(dfu[b])), repr(float(dfv[b])), repr(float(dfa[b]))]
            fh.write("\t".join(row) + "\n")
