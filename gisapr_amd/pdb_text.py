"""Text PDB input and the rigid-body pose transform that feeds the splat.

Mirrors pdb_text.py (read_pdb :78-120, transform_coordinates :123-142,
transform_chain_coordinates :145-150) and the zip release's
transformed_density_atoms (first-model transform, all-model density rows,
PDB 8.3 coordinate round trip).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass
from functools import lru_cache

import numpy as np

_ELEMENTS = set((
    "H HE LI BE B C N O F NE NA MG AL SI P S CL AR K CA SC TI V CR MN FE "
    "CO NI CU ZN GA GE AS SE BR KR RB SR Y ZR NB MO TC RU RH PD AG CD IN "
    "SN SB TE I XE CS BA LA CE PR ND PM SM EU GD TB DY HO ER TM YB LU HF "
    "TA W RE OS IR PT AU HG TL PB BI PO AT RN FR RA AC TH PA U NP PU AM "
    "CM BK CF ES FM MD NO LR RF DB SG BH HS MT DS RG CN NH FL MC LV TS OG D").split())


def infer_element(line: str) -> str:
    """Element of an ATOM/HETATM line: columns 77-78, else the aligned atom name."""
    explicit = line[76:78].strip().upper()
    if explicit in _ELEMENTS:
        return explicit
    name = line[12:16].ljust(4).upper()
    letters = "".join(ch for ch in name if ch.isalpha())
    if not letters:
        return ""
    if name[0].isspace() or name[0].isdigit():
        return letters[0] if letters[0] in _ELEMENTS else ""
    if line[:6] == "ATOM  " and letters[0] in "CNOHPS":
        return letters[0]
    if letters[:2] in _ELEMENTS:
        return letters[:2]
    return letters[0] if letters[0] in _ELEMENTS else ""


@dataclass
class PDBText:
    lines: list
    coords: np.ndarray          # [A, 3] float32
    elements: np.ndarray        # [A] str
    chains: np.ndarray          # [A] str
    atom_line_indices: np.ndarray

    def write(self, output, coords=None) -> None:
        """Every original line, with the x/y/z columns (31-54) of this object's
        atom records replaced by ``coords`` in 8.3 format when given
        (pdb_text.py:59-75)."""
        lines = list(self.lines)
        if coords is not None:
            xyz = np.asarray(coords)
            if xyz.shape != self.coords.shape:
                raise ValueError("Replacement coordinates must have shape %s" % (self.coords.shape,))
            if not np.all(np.isfinite(xyz)):
                raise ValueError("PDB coordinates must be finite")
            for li, row in zip(self.atom_line_indices, xyz):
                ln = lines[int(li)]
                lines[int(li)] = ln[:30] + "".join(coordinate_fields(row)) + ln[54:]
        with open(output, "w", encoding="utf-8", errors="surrogateescape", newline="") as fh:
            fh.writelines(lines)


def read_pdb(path, include_hetatm=True, first_model_only=True) -> PDBText:
    """xyz, element and chain of ATOM (and HETATM) records, in file order."""
    with open(path, "r", encoding="utf-8", errors="surrogateescape", newline="") as fh:
        lines = fh.readlines()
    first_model = next((i for i, ln in enumerate(lines) if ln[:6].strip() == "MODEL"), None)
    active = first_model is None
    kinds = ("ATOM", "HETATM") if include_hetatm else ("ATOM",)
    coords, elements, chains, idx = [], [], [], []
    for i, ln in enumerate(lines):
        rec = ln[:6].strip()
        if rec == "MODEL":
            if first_model_only and i != first_model:
                break
            active = True
            continue
        if rec == "ENDMDL" and first_model is not None:
            if first_model_only:
                break
            active = False
            continue
        if not active or rec not in kinds:
            continue
        try:
            xyz = [float(ln[s:s + 8]) for s in (30, 38, 46)]
        except ValueError as exc:
            raise ValueError("Invalid PDB coordinates in %s, line %d" % (os.fspath(path), i + 1)) from exc
        if not all(math.isfinite(v) for v in xyz):
            raise ValueError("Non-finite PDB coordinates in %s, line %d" % (os.fspath(path), i + 1))
        coords.append(xyz)
        elements.append(infer_element(ln))
        chains.append(ln[21:22] or " ")
        idx.append(i)
    return PDBText(lines, np.asarray(coords, np.float32).reshape(-1, 3), np.asarray(elements, "U2"),
                   np.asarray(chains, "U1"), np.asarray(idx, np.int64))


@lru_cache(maxsize=16)
def _read_pdb_cached(path, mtime, include_hetatm, first_model_only):
    return read_pdb(path, include_hetatm, first_model_only)


def read_pdb_cached(path, include_hetatm=True, first_model_only=True) -> PDBText:
    return _read_pdb_cached(os.fspath(path), os.path.getmtime(path), include_hetatm, first_model_only)


def rotvec_matrix(rotvec_deg) -> np.ndarray:
    """Rotation matrix of a rotation vector given in degrees (Rodrigues)."""
    v = np.deg2rad(np.asarray(rotvec_deg, np.float64))
    th = float(np.linalg.norm(v))
    if th == 0.0:
        return np.eye(3)
    k = v / th
    K = np.array([[0.0, -k[2], k[1]], [k[2], 0.0, -k[0]], [-k[1], k[0], 0.0]])
    return np.eye(3) * math.cos(th) + (1.0 - math.cos(th)) * np.outer(k, k) + math.sin(th) * K


def transform_coordinates(coords, pose, pivot=None) -> np.ndarray:
    """(coords - pivot) @ R(rotvec deg) + pivot + translation (pdb_text.py:123-142)."""
    coords = np.asarray(coords)
    pose = np.asarray(pose, np.float64)
    if coords.ndim != 2 or coords.shape[1] != 3 or not len(coords):
        raise ValueError("A nonempty (N, 3) coordinate array is required")
    if pose.shape != (6,) or not np.all(np.isfinite(pose)):
        raise ValueError("Pose must contain six finite rotation/translation values")
    center = np.mean(coords, axis=0) if pivot is None else np.asarray(pivot, np.float64)
    if np.shape(center) != (3,) or not np.all(np.isfinite(center)):
        raise ValueError("Pivot must contain three finite coordinates")
    return (coords - center) @ rotvec_matrix(pose[:3]) + center + pose[3:]


def transform_chain_coordinates(pdb: PDBText, chain_id, pose, pivot=None) -> np.ndarray:
    sel = pdb.chains == str(chain_id)
    if not np.any(sel):
        raise ValueError("Chain %r has no atoms in the first PDB model" % chain_id)
    return transform_coordinates(pdb.coords[sel], pose, pivot=pivot)


def coordinate_fields(xyz):
    """PDB 8.3 text fields of one coordinate triple."""
    fields = [f"{float(v):8.3f}" for v in xyz]
    if any(len(f) != 8 for f in fields):
        raise ValueError("Coordinate exceeds the PDB 8.3 field width")
    return fields


def quantize_8_3(xyz) -> np.ndarray:
    """float(f"{x:8.3f}") of every coordinate (the write/re-read round trip), float32.

    Vectorised: k = round-half-even(1000 x) on the exact binary value of x,
    then k / 1000 (correctly rounded, as float() of the 8.3 text is).  The
    few values whose 1000 x lies within 1e-6 of a tie are decided exactly
    from the formatted text."""
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
        for r in xyz.reshape(-1, 3):
            coordinate_fields(r)  # raises with the reference's message
    return out.astype(np.float32).reshape(-1, 3)


def transformed_density_atoms(path, chain_id, pose, include_hetatm=False):
    """Density atoms after moving ``chain_id`` by ``pose``: the first-model
    transform (pivot over first-model rows incl. HETATM), written to 8.3 text
    and re-read over all models."""
    first = read_pdb_cached(path)
    dens = read_pdb_cached(path, include_hetatm=include_hetatm, first_model_only=False)
    sel = first.chains == str(chain_id)
    moved = transform_chain_coordinates(first, chain_id, pose)
    if not np.isfinite(moved).all():
        raise ValueError("PDB coordinates must be finite")
    rows = {int(li): p for p, li in enumerate(dens.atom_line_indices)}
    coords = dens.coords.copy()
    q = quantize_8_3(moved)
    for li, xyz in zip(first.atom_line_indices[sel], q):
        p = rows.get(int(li))
        if p is not None:
            coords[p] = xyz
    return coords, dens.elements


def move_chain(coords, chains, chain_id, pose, quantize=True) -> np.ndarray:
    """In-memory form of transformed_density_atoms for an already-parsed model."""
    coords = np.asarray(coords, np.float32).copy()
    sel = np.asarray(chains) == str(chain_id)
    moved = transform_coordinates(coords[sel], pose)
    coords[sel] = quantize_8_3(moved) if quantize else moved.astype(np.float32)
    return coords


def write_transformed_chain(path, chain_id, pose, output, pivot=None) -> np.ndarray:
    """The whole PDB with only the first-model chain ``chain_id`` moved by
    ``pose``; every other line is kept byte for byte (pdb_text.py:153-162).
    Returns the chain's new coordinates."""
    pdb = read_pdb(path)
    sel = pdb.chains == str(chain_id)
    moved = transform_chain_coordinates(pdb, chain_id, pose, pivot=pivot)
    PDBText(pdb.lines, pdb.coords[sel], pdb.elements[sel], pdb.chains[sel], pdb.atom_line_indices[sel]).write(
        output, moved)
    return moved


def split_pdb_by_chain(input_pdb, chain_id, output_chain_file, output_rest_file) -> None:
    """The first model's atom records of chain ``chain_id``, and all the others,
    verbatim into two files, each ending with END (pdb_text.py:165-183)."""
    pdb = read_pdb(input_pdb)
    sel = pdb.chains == str(chain_id)
    if not np.any(sel):
        raise ValueError("Chain %r has no atoms in the first PDB model" % chain_id)
    for output, mask in ((output_chain_file, sel), (output_rest_file, ~sel)):
        with open(output, "w", encoding="utf-8", errors="surrogateescape", newline="") as fh:
            for li in pdb.atom_line_indices[mask]:
                ln = pdb.lines[int(li)]
                fh.write(ln if ln.endswith(("\n", "\r")) else ln + "\n")
            fh.write("END\n")
