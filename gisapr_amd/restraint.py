"""Geometric restraint of a moving chain against the fixed main body, on the
HIP splat (func_check_boundary_for_testing_v8.py:62-126).

The main body (every atom outside the chain) is splatted once at the
restraint box; its voxels above ``threshold`` form a mask, whose exact
Euclidean distance transform stays resident on the device.  A pose moves the
chain (rotation vector in degrees about the chain's centre, then a shift in
A, float32 row vectors as the reference does), splats it, and returns

  * the number of voxels above threshold in both the chain and the body, and
  * when there is no overlap, the smallest distance (A) from a chain voxel to
    the body: min over chain voxels of the body's distance transform; when the
    chain has no voxel above threshold, the distance from the point
    (-1, 0, 0) (in voxels) to the body -- the value scipy's transform of an
    all-background grid gives, kept by the reference (:118-125).

Same signature, defaults and return values as the reference class; the
distance transform and the reductions are kernels of this package
(gis_edt_squared, gis_restraint_eval) instead of SciPy and torch.
"""
from __future__ import annotations

import numpy as np

from .density import atoms_to_atomic_numbers_and_masses, density_from_coords

OUT_OF_BOUNDARY = False
IN_BOUNDARY = True


def read_pdb_split_chain_and_rest(pdb_filename, target_chain_id, do_include_HETATM=False):
    """(chain coords, chain elements, rest coords, rest elements) over every
    model (func_check_boundary_for_testing_v8.py:24-30)."""
    from .pdb_text import read_pdb
    pdb = read_pdb(pdb_filename, include_hetatm=do_include_HETATM, first_model_only=False)
    sel = pdb.chain_mask(target_chain_id)
    if not sel.any():
        raise ValueError(f"Cannot find chain {target_chain_id}")
    return pdb.coords[sel], pdb.elements[sel], pdb.coords[~sel], pdb.elements[~sel]


def rotate_coords_rotvec(coordinates, rot, tilt, psi, xshift, yshift, zshift, pivot=None):
    """float32 row vectors moved by the rotation vector (rot, tilt, psi) in
    degrees about ``pivot`` (default: their mean), then shifted (A)
    (func_check_boundary_for_testing_v8.py:33-45)."""
    from .pdb_text import rotvec_matrix
    xyz = np.asarray(coordinates, np.float32)
    if xyz.shape[0] == 0:
        return xyz.copy()
    centre = np.mean(xyz, axis=0) if pivot is None else np.asarray(pivot, np.float32)
    if centre.shape != (3,) or not np.isfinite(centre).all():
        raise ValueError("pivot must contain three finite coordinates in Angstrom")
    angles = np.asarray([rot, tilt, psi], np.float32).astype(np.float64)
    rmat = rotvec_matrix(angles).astype(np.float32)
    shift = np.asarray([xshift, yshift, zshift], np.float32)
    return ((xyz - centre) @ rmat + centre + shift).astype(np.float32, copy=False)


class GeometryRestraint:
    """Resident main-body mask and distance transform for one chain of a model."""

    def __init__(self, pdb_path, chain_id, device="cpu", boxsize=256, apix=1.5, threshold=1.0,
                 include_hetatm=False):
        import torch
        from . import kernels
        from ._lib import check, lib, ptr, stream_of
        self._torch, self._check, self._lib, self._ptr, self._stream = torch, check, lib, ptr, stream_of
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise ValueError(f"GeometryRestraint runs on HIP devices only: pass device='cuda' (got {self.device})")
        self.boxsize = int(boxsize)
        self.apix = float(apix)
        self.threshold = float(threshold)
        chain_xyz, chain_el, rest_xyz, rest_el = read_pdb_split_chain_and_rest(pdb_path, chain_id, include_hetatm)
        self.chain_coords = chain_xyz
        numbers, _ = atoms_to_atomic_numbers_and_masses(chain_el)
        self.chain_numbers = torch.as_tensor(numbers, dtype=torch.float32, device=self.device)
        rest_numbers, _ = atoms_to_atomic_numbers_and_masses(rest_el)
        body = density_from_coords(rest_xyz, rest_numbers, self.boxsize, self.apix, 2.0 * self.apix,
                                   device=self.device)
        self.mainbody_mask = self._mask(body)
        del body
        self.has_mainbody = bool(self.mainbody_mask.any())
        self.mainbody_d2 = None
        self._empty_chain_distance = None
        if self.has_mainbody:
            self.mainbody_d2 = torch.empty((self.boxsize,) * 3, dtype=torch.float32, device=self.device)
            check(lib().gis_edt_squared(ptr(self.mainbody_mask), self.boxsize, ptr(self.mainbody_d2),
                                        stream_of(self.device)), "gis_edt_squared")
        self._out = torch.zeros(4, dtype=torch.int64, device=self.device)  # overlap, min_d2 bits, any

    def _mask(self, vol):
        m = self._torch.empty(vol.shape, dtype=self._torch.uint8, device=self.device)
        self._check(self._lib().gis_threshold_mask(self._ptr(vol), vol.numel(), self.threshold, self._ptr(m),
                                                   self._stream(self.device)), "gis_threshold_mask")
        return m

    @property
    def mainbody_distance(self):
        """The body's distance transform in voxels (float64, as SciPy returns it)."""
        if self.mainbody_d2 is None:
            return None
        return self._torch.sqrt(self.mainbody_d2.double())

    def evaluate(self, pose, pivot=None):
        """(overlapped voxel count, minimum separation in A) of the chain moved
        by ``pose`` = [rot, tilt, psi, xshift, yshift, zshift]."""
        torch = self._torch
        pose = np.asarray(pose, np.float64)
        if pose.shape != (6,) or not np.isfinite(pose).all():
            raise ValueError("pose must contain six finite rotation/translation values")
        if not self.has_mainbody:
            return 0, 0.0
        xyz = rotate_coords_rotvec(self.chain_coords, *pose, pivot=pivot)
        vol = density_from_coords(xyz, self.chain_numbers, self.boxsize, self.apix, 2.0 * self.apix,
                                  device=self.device)
        o = self._out
        base = o.data_ptr()
        import ctypes as C
        self._check(self._lib().gis_restraint_eval(self._ptr(vol), self.threshold, self._ptr(self.mainbody_mask),
                                                   self._ptr(self.mainbody_d2), vol.numel(), C.c_void_p(base),
                                                   C.c_void_p(base + 8), C.c_void_p(base + 16),
                                                   self._stream(self.device)), "gis_restraint_eval")
        overlap, d2_bits, any_chain = (int(v) for v in o[:3].cpu().tolist())
        if overlap:
            return overlap, 0.0
        if any_chain:
            d2 = float(np.array([d2_bits & 0xFFFFFFFF], np.uint32).view(np.float32)[0])
            return 0, float(np.sqrt(np.float64(d2))) * self.apix
        if self._empty_chain_distance is None:
            zz, yy, xx = (t.double() for t in torch.nonzero(self.mainbody_mask, as_tuple=True))
            self._empty_chain_distance = float(torch.sqrt(((zz + 1) ** 2 + yy ** 2 + xx ** 2).min())) * self.apix
        return 0, self._empty_chain_distance
