"""RELION STAR loops: the direction table and the particle table.

Selection rules the v606 readers apply (v606:75-316), kept so the same files
give the same rows:
  * a ``loop_`` line opens a loop; the ``_label`` lines after it (blank and
    ``#`` comment lines skipped) name its columns, each by its ``#n`` suffix
    (1-based) or, without one, by its position among the labels;
  * the loop's data rows are the following lines that are not blank or
    comments, up to the next ``loop_`` / ``data_`` / ``_label`` line once a
    row has been read;
  * the first loop whose labels include every required label is used;
  * a row with too few columns for the needed labels is skipped.

Every line is classified once (comment, loop_, data_, label, row); the
readers take the selected loop and pull typed columns out of it.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, Iterator, List, Optional, Sequence

import numpy as np


@dataclass
class StarLoop:
    lines: List[str]
    labels: Dict[str, int]       # label -> 0-based column
    data_line_indices: List[int]  # 0-based line number of each data row
    data_start: int              # line number of the first data row, -1 if none

    def tokens(self, line_index: int) -> List[str]:
        return self.lines[line_index].split()


def _label_column(tokens: Sequence[str], position: int) -> int:
    """Column of a label line: its ``#n`` suffix (1-based) or its position."""
    for tok in tokens[1:]:
        if tok[:1] == "#" and tok[1:].isdigit():
            return int(tok[1:]) - 1
    return position


def _kind(line: str) -> str:
    s = line.strip()
    if not s or s[0] == "#":
        return "skip"
    low = s.lower()
    if low == "loop_":
        return "loop"
    if low.startswith("data_"):
        return "block"
    if s[0] == "_":
        return "label"
    return "row"


def _loop_headers(kinds: List[str], lines: List[str]) -> Iterator[tuple]:
    """(labels, index of the first line after the label block) of every loop."""
    n = len(lines)
    k = 0
    while k < n:
        if kinds[k] != "loop":
            k += 1
            continue
        labels: Dict[str, int] = {}
        j = k + 1
        while j < n and kinds[j] in ("label", "skip"):
            if kinds[j] == "label":
                tok = lines[j].split()
                labels[tok[0]] = _label_column(tok, len(labels))
            j += 1
        yield labels, j
        k = j


def _data_rows(kinds: List[str], start: int) -> List[int]:
    """Data rows from ``start``: a loop_ / data_ / label line ends them once a
    row has been read (before the first row it is passed over)."""
    rows: List[int] = []
    for k in range(start, len(kinds)):
        kind = kinds[k]
        if kind == "row":
            rows.append(k)
        elif kind != "skip" and rows:
            break
    return rows


def find_star_loop(lines: List[str], required_labels: Sequence[str]) -> StarLoop:
    """The first loop naming every required label (ValueError if there is
    none, or if it has no data rows)."""
    need = set(required_labels)
    kinds = [_kind(ln) for ln in lines]
    for labels, after in _loop_headers(kinds, lines):
        if need.issubset(labels):
            rows = _data_rows(kinds, after)
            if not rows:
                raise ValueError("STAR loop with " + ", ".join(required_labels) + " has no data rows")
            return StarLoop(lines, labels, rows, rows[0])
    raise ValueError("no STAR loop with " + ", ".join(required_labels))


def read_text_lines(filename: str) -> List[str]:
    with open(filename, "r", encoding="utf-8", errors="replace") as fh:
        return fh.readlines()


def _columns(loop: StarLoop, required: Sequence[str], optional: Sequence[str] = ()):
    """(line index, {label: token}) for each data row holding every required
    column; optional labels map to None when absent from the row."""
    req = {lab: loop.labels[lab] for lab in required}
    opt = {lab: loop.labels.get(lab, -1) for lab in optional}
    need = max(req.values())
    for li in loop.data_line_indices:
        tok = loop.tokens(li)
        if len(tok) <= need:
            continue
        row = {lab: tok[c] for lab, c in req.items()}
        row.update({lab: (tok[c] if 0 <= c < len(tok) else None) for lab, c in opt.items()})
        yield li, row


@dataclass
class ModelAngles:
    rot: List[float]
    tilt: List[float]
    psi: List[float]
    rotation_matrices_t: np.ndarray  # [M, 3, 3] float32, Euler(rot, tilt, 0)^T


def read_model_angles(angle_star: str) -> ModelAngles:
    """The direction table (v606:190-225); projections use rot and tilt only."""
    from .host import rotation_matrices_t
    loop = find_star_loop(read_text_lines(angle_star), ["_rlnAngleRot", "_rlnAngleTilt"])
    rows = list(_columns(loop, ["_rlnAngleRot", "_rlnAngleTilt"], ["_rlnAnglePsi"]))
    if not rows:
        raise ValueError(f"{angle_star}: the direction table has no complete rows")
    rot = [float(r["_rlnAngleRot"]) for _, r in rows]
    tilt = [float(r["_rlnAngleTilt"]) for _, r in rows]
    psi = [float(r["_rlnAnglePsi"]) if r["_rlnAnglePsi"] is not None else 0.0 for _, r in rows]
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


_CTF_LABELS = ("_rlnDefocusU", "_rlnDefocusV", "_rlnDefocusAngle")


def read_particle_table(particle_star: str) -> ParticleTable:
    """Particle rows with their priors and, when the loop names all three
    defocus columns, their defocus (v606:249-297)."""
    lines = read_text_lines(particle_star)
    loop = find_star_loop(lines, ["_rlnImageName", "_rlnAngleRot", "_rlnAngleTilt"])
    has_ctf = all(lab in loop.labels for lab in _CTF_LABELS)
    required = ["_rlnImageName", "_rlnAngleRot", "_rlnAngleTilt", *(_CTF_LABELS if has_ctf else ())]
    records: List[ParticleRecord] = []
    for li, r in _columns(loop, required, ["_rlnAnglePsi"]):
        rec = ParticleRecord(li, r["_rlnImageName"], float(r["_rlnAngleRot"]), float(r["_rlnAngleTilt"]),
                             float(r["_rlnAnglePsi"]) if r["_rlnAnglePsi"] is not None else 0.0)
        if has_ctf:
            rec.defocus_u, rec.defocus_v, rec.defocus_angle = (float(r[lab]) for lab in _CTF_LABELS)
        records.append(rec)
    if not records:
        raise ValueError(f"{particle_star}: the particle table has no complete rows")
    return ParticleTable(lines, loop, records, {r.line_index: r for r in records}, has_ctf)


def select_particle_records(table: ParticleTable, start_line: int, end_line: int) -> List[ParticleRecord]:
    """The rows whose line index lies in [start, end): -1 leaves a side open,
    and the range never starts before the loop's first data row
    (--startLine / --endLine, v606:300-309)."""
    first = table.loop.data_start
    lo = first if start_line < 0 else max(start_line, first)
    hi = len(table.lines) if end_line < 0 else max(min(end_line, len(table.lines)), first)
    return [r for r in table.records if lo <= r.line_index < hi]


def write_direction_star(path, rot, tilt, psi: Optional[Sequence[float]] = None):
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
        cols += list(_CTF_LABELS)
    with open(path, "w") as fh:
        fh.write("# relion 30001\n\ndata_particles\n\nloop_\n")
        for i, c in enumerate(cols):
            fh.write(f"{c} #{i + 1}\n")
        for b in range(len(rot)):
            row = [f"{b + 1}@{stack_name}", repr(float(rot[b])), repr(float(tilt[b])), repr(float(psi[b]))]
            if dfu is not None:
                row += [repr(float(dfu[b])), repr(float(dfv[b])), repr(float(dfa[b]))]
            fh.write("\t".join(row) + "\n")
