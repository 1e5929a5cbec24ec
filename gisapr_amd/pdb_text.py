"""Atom records of a text PDB and the rigid-body pose that moves one chain.

The candidate path needs three things from a PDB file: the coordinates,
element symbols and chain letters of its atom records (for the splat), and
the move of one chain by a pose -- rotation vector in degrees about the
chain's own centre, then a translation in A -- followed by the fixed-column
8.3 text round trip the reference does when it writes the moved model and
reads it back (pdb_text.py:123-150; Z3 pdb_text.py:179-198).

Record selection (the PDB format): columns 1-6 name the record; atom
coordinates sit in columns 31-38, 39-46, 47-54; the chain letter in column
22; the element symbol in columns 77-78.  When the file has MODEL records,
only atoms inside MODEL ... ENDMDL blocks count, and the pose pivot uses the
first block only.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass
from functools import lru_cache
from typing import List, Optional, Tuple

import numpy as np

# symbols accepted from columns 77-78 or derived from the atom name
_KNOWN = frozenset("""
H HE LI BE B C N O F NE NA MG AL SI P S CL AR K CA SC TI V CR MN FE CO NI CU ZN GA GE AS SE BR KR RB SR Y ZR
NB MO TC RU RH PD AG CD IN SN SB TE I XE CS BA LA CE PR ND PM SM EU GD TB DY HO ER TM YB LU HF TA W RE OS IR
PT AU HG TL PB BI PO AT RN FR RA AC TH PA U NP PU AM CM BK CF ES FM MD NO LR RF DB SG BH HS MT DS RG CN NH FL
MC LV TS OG D""".split())

# one-letter elements that an ATOM (polymer) record's atom name starts with
_POLYMER_FIRST = frozenset("CNOHPS")


def infer_element(record: str) -> str:
    """Element symbol of an ATOM / HETATM line ("" when none can be read).

    Columns 77-78 win when they hold a known symbol.  Otherwise the atom name
    (columns 13-16) decides: a name whose first column is blank or a digit is
    a one-letter element in column 14; an ATOM record's name starting with
    C, N, O, H, P or S is that element (CA in a protein is carbon alpha); any
    other name is the two-letter symbol it starts with when that is known,
    else its first letter."""
    sym = record[76:78].strip().upper()
    if sym in _KNOWN:
        return sym
    name = record[12:16].ljust(4).upper()
    letters = [ch for ch in name if ch.isalpha()]
    if not letters:
        return ""
    first = letters[0]
    one = first if first in _KNOWN else ""
    if name[0].isspace() or name[0].isdigit():
        return one
    if record.startswith("ATOM  ") and first in _POLYMER_FIRST:
        return first
    two = "".join(letters[:2])
    return two if len(two) == 2 and two in _KNOWN else one


@dataclass
class AtomRecords:
    """Atom records in file order: [A, 3] float32 coordinates, element and
    chain per atom, and the 0-based line number of each record."""
    lines: List[str]
    coords: np.ndarray
    elements: np.ndarray
    chains: np.ndarray
    line_numbers: np.ndarray

    def chain_mask(self, chain_id) -> np.ndarray:
        return self.chains == str(chain_id)


def _model_blocks(records: List[str]) -> Optional[List[Tuple[int, int]]]:
    """[start, end) line ranges of the MODEL ... ENDMDL blocks, or None when
    the file has no MODEL record (then every line counts)."""
    blocks, start = [], None
    for n, line in enumerate(records):
        tag = line[:6].strip()
        if tag == "MODEL":
            if start is not None:  # an unterminated block ends at the next MODEL
                blocks.append((start, n))
            start = n + 1
        elif tag == "ENDMDL" and start is not None:
            blocks.append((start, n))
            start = None
    if start is not None:
        blocks.append((start, len(records)))
    if not blocks and not any(line[:6].strip() == "MODEL" for line in records):
        return None
    return blocks


def _coordinates(line: str, where: str) -> List[float]:
    try:
        xyz = [float(line[30:38]), float(line[38:46]), float(line[46:54])]
    except ValueError as exc:
        raise ValueError(f"unreadable atom coordinates at {where}") from exc
    if not all(math.isfinite(v) for v in xyz):
        raise ValueError(f"non-finite atom coordinates at {where}")
    return xyz


def read_pdb(path, include_hetatm=True, first_model_only=True) -> AtomRecords:
    """ATOM (and, with ``include_hetatm``, HETATM) records of a text PDB; only
    the first MODEL block with ``first_model_only`` (the pose pivot's atoms,
    pdb_text.py:78-120), every block otherwise (the density's atoms)."""
    with open(path, "r", encoding="utf-8", errors="surrogateescape", newline="") as fh:
        records = fh.readlines()
    blocks = _model_blocks(records)
    if blocks is None:
        spans = [(0, len(records))]
    else:
        spans = blocks[:1] if first_model_only else blocks
    wanted = ("ATOM", "HETATM") if include_hetatm else ("ATOM",)
    xyz, elem, chain, where = [], [], [], []
    for lo, hi in spans:
        for n in range(lo, hi):
            line = records[n]
            if line[:6].strip() not in wanted:
                continue
            xyz.append(_coordinates(line, f"{os.fspath(path)}:{n + 1}"))
            elem.append(infer_element(line))
            chain.append(line[21:22] or " ")
            where.append(n)
    return AtomRecords(records, np.asarray(xyz, np.float32).reshape(-1, 3), np.asarray(elem, "U2"),
                       np.asarray(chain, "U1"), np.asarray(where, np.int64))


@lru_cache(maxsize=16)
def _read_cached(path, mtime_ns, include_hetatm, first_model_only):
    return read_pdb(path, include_hetatm, first_model_only)


def read_pdb_cached(path, include_hetatm=True, first_model_only=True) -> AtomRecords:
    """read_pdb, reused while the file is unchanged (the optimizers re-read one
    model for every candidate)."""
    return _read_cached(os.fspath(path), os.stat(path).st_mtime_ns, include_hetatm, first_model_only)


def rotvec_matrix(rotvec_deg) -> np.ndarray:
    """3x3 matrix of the rotation by |v| degrees about v (Rodrigues),
    acting on column vectors (scipy's Rotation.from_rotvec(...).as_matrix())."""
    v = np.radians(np.asarray(rotvec_deg, np.float64).reshape(3))
    angle = float(np.sqrt(v @ v))
    if angle == 0.0:
        return np.eye(3)
    a = v / angle
    cross = np.array([[0.0, -a[2], a[1]], [a[2], 0.0, -a[0]], [-a[1], a[0], 0.0]])
    return np.eye(3) + math.sin(angle) * cross + (1.0 - math.cos(angle)) * (cross @ cross)


def transform_coordinates(coords, pose, pivot=None) -> np.ndarray:
    """Rigid move of row-vector coordinates: (x - pivot) R + pivot + t, with R
    the rotation matrix of pose[:3] (degrees) and t = pose[3:] (A); the pivot
    defaults to the coordinates' mean (pdb_text.py:123-142: row vectors times
    the matrix, i.e. the inverse rotation of each point)."""
    xyz = np.asarray(coords)
    if xyz.ndim != 2 or xyz.shape[1] != 3 or xyz.shape[0] == 0:
        raise ValueError("coordinates must be a nonempty (N, 3) array")
    p = np.asarray(pose, np.float64)
    if p.shape != (6,) or not np.isfinite(p).all():
        raise ValueError("pose must be six finite numbers: rotation vector (deg) and translation (A)")
    centre = xyz.mean(axis=0) if pivot is None else np.asarray(pivot, np.float64)
    if centre.shape != (3,) or not np.isfinite(centre).all():
        raise ValueError("pivot must be three finite numbers")
    return (xyz - centre) @ rotvec_matrix(p[:3]) + centre + p[3:]


def transform_chain_coordinates(pdb: AtomRecords, chain_id, pose, pivot=None) -> np.ndarray:
    """The moved coordinates of one chain of a parsed model (pdb_text.py:145-150)."""
    mask = pdb.chain_mask(chain_id)
    if not mask.any():
        raise ValueError(f"chain {chain_id!r} has no atom records")
    return transform_coordinates(pdb.coords[mask], pose, pivot=pivot)


def quantize_8_3(xyz) -> np.ndarray:
    """float(f"{x:8.3f}") of every coordinate -- the PDB write / re-read round
    trip -- as float32 [N, 3].

    Vectorised: k = round-half-even(1000 x) on the exact binary value of x,
    then k / 1000 (correctly rounded, as float() of the 8.3 text is).  The
    few values whose 1000 x lies within 1e-6 of a tie are decided exactly
    from the formatted text.  Values that do not fit the 8-character field
    raise, as writing them would."""
    xyz = np.asarray(xyz, np.float64)
    if xyz.size and not np.all(np.isfinite(xyz)):
        raise ValueError("PDB coordinates must be finite")
    t = xyz * 1000.0
    k = np.rint(t)
    frac = np.abs(t - np.trunc(t))
    near = np.abs(frac - 0.5) < 1e-6
    if np.any(near):
        idx = np.nonzero(near)
        k[idx] = [round(float(f"{v:.3f}") * 1000.0) for v in xyz[idx]]
    out = k / 1000.0
    if out.size and (out.max() >= 9999.9995 or out.min() <= -999.9995):
        raise ValueError("a coordinate exceeds the PDB 8.3 field width")
    return out.astype(np.float32).reshape(-1, 3)


def transformed_density_atoms(path, chain_id, pose, include_hetatm=False):
    """Density atoms (all models) after moving ``chain_id`` of the first model
    by ``pose`` about its first-model centre (every first-model record of the
    chain, HETATM included, sets the pivot) and the 8.3 round trip
    (Z3 pdb_text.py:179-198).  Returns (coords [A, 3] float32, elements)."""
    first = read_pdb_cached(path)
    dens = read_pdb_cached(path, include_hetatm=include_hetatm, first_model_only=False)
    mask = first.chain_mask(chain_id)
    moved = quantize_8_3(transform_chain_coordinates(first, chain_id, pose))
    slot = {int(n): k for k, n in enumerate(dens.line_numbers)}
    coords = dens.coords.copy()
    for n, xyz in zip(first.line_numbers[mask], moved):
        k = slot.get(int(n))
        if k is not None:
            coords[k] = xyz
    return coords, dens.elements


def move_chain(coords, chains, chain_id, pose, quantize=True) -> np.ndarray:
    """transformed_density_atoms for an already-parsed model: the rows of
    ``chain_id`` moved by ``pose`` about their mean (and 8.3-rounded)."""
    coords = np.asarray(coords, np.float32).copy()
    mask = np.asarray(chains) == str(chain_id)
    moved = transform_coordinates(coords[mask], pose)
    coords[mask] = quantize_8_3(moved) if quantize else moved.astype(np.float32)
    return coords
