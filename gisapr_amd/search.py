"""Drop-in replacement of the v606 search module on the HIP engine.

Keeps the call signatures and the output contract of
test1_with_isspa_weight_varingKK_search_translation_also_v606_torch_optimized_standalone.py
(``v606`` below) and its zip-release form:

* ``create_search_parser()`` -- the v606 argparse surface (v606:1524-1568 plus
  the zip release's --verbose / cache options)
* ``run_search(args, *, model_volume=None, session_cache=None) -> dict``
  (v606:1594-1825; zip release: tensor model input, returned metrics)
* ``choose_device``, ``chunks``, ``main`` (the --worker-connect dispatch)
* the helpers the batched engine resolves from ``run_search.__globals__``:
  PackedFourierLayout, read_model_angles, read_particle_table,
  select_particle_records, build_local_orientation_entries,
  generate_soft_mask, read_fsc_map, IntegratedProjector,
  psi_sampling_for_record, PsiGridBank, ParticleFourierPreparer,
  PackedCTFTorch, PackedCTFNumPy, ParticleMRCReader, CorrelationWorkspace,
  correlate_templates_and_particle, make_result_records,
  calculate_angular_distance, euler_angles_to_matrix, and the grid_sample
  projector's pieces apply_sinc2_filter_inplace_3d, build_projection_base_grid,
  rotated_grid_to_grid_sample, ProjectionWhitening.

All array work runs in the HIP kernels (libgisapr_hip.so); there is no CPU
path -- on a host without a HIP device ``choose_device`` raises.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import host, kernels, mrc, orientation_bank
from .engine import SearchEngine
from .projection import (ProjectionWhitening, apply_sinc2_filter_inplace_3d,  # noqa: F401
                         build_projection_base_grid, rotated_grid_to_grid_sample)
from .host import (calculate_angular_distance, euler_angles_to_matrix, expand_radial_values,  # noqa: F401
                   generate_soft_mask, psi_sampling_for_record)
from .star import (ModelAngles, ParticleRecord, ParticleTable, StarLoop, find_star_loop,  # noqa: F401
                   read_model_angles, read_particle_table, read_text_lines, select_particle_records)
from .star import _label_column as _parse_label_index  # noqa: F401  (v606:86-92)

torch.set_grad_enabled(False)


# ---------------------------------------------------------------------------
# Orientation helpers
# ---------------------------------------------------------------------------

def build_local_orientation_entries(particle_records, model_rot, model_tilt, local_range) -> List[List[int]]:
    """[[line_index, model indices...], ...] (v606:351-378)."""
    rot = np.asarray([r.rot for r in particle_records], np.float64)
    tilt = np.asarray([r.tilt for r in particle_records], np.float64)
    local = host.local_orientation_map(rot, tilt, model_rot, model_tilt, float(local_range))
    return [[r.line_index] + l for r, l in zip(particle_records, local)]


def clip_value(a: float, b: float, c: float) -> float:
    """a clamped to [b, c] as float (v606:314-319)."""
    return float(min(max(float(a), float(b)), float(c)))


def dot_product(v1: Sequence[float], v2: Sequence[float]) -> float:
    """Sum of products, or -9999999.0 for lengths that differ (v606:322-328)."""
    if len(v1) != len(v2):
        return -9999999.0
    total = 0.0
    for a, b in zip(v1, v2):
        total += float(a) * float(b)
    return total


def _local_cache_valid(entries, particle_records, n_models: int) -> bool:
    """A cached local map matches the selected records row for row and holds
    valid direction indices (v606:381-393)."""
    if not isinstance(entries, list) or len(entries) != len(particle_records):
        return False
    for entry, rec in zip(entries, particle_records):
        if not isinstance(entry, list) or not entry or entry[0] != rec.line_index:
            return False
        if any(not isinstance(i, (int, np.integer)) or i < 0 or i >= n_models for i in entry[1:]):
            return False
    return True


class _PlainDataUnpickler:
    """Reads a pickle of plain lists and ints only: any other object (a class
    or function reference, which is how a pickle runs code) is refused."""

    def __init__(self, fh):
        import pickle

        class _Restricted(pickle.Unpickler):
            def find_class(self, module, name):
                raise pickle.UnpicklingError(f"refused object {module}.{name} in a local-orientation cache")

        self._u = _Restricted(fh)

    def load(self):
        return self._u.load()


def load_or_build_local_orientation_map(particle_star: str, particle_records, model_angles: ModelAngles,
                                        local_range: float) -> Dict[int, List[int]]:
    """{line_index: direction indices} with the v605/v606 cache file
    ``<star>_LocalSearch_Angle_<int(range)>_data.pkl`` (v606:396-438): a valid
    cache is reused, otherwise the map is built and the cache rewritten
    atomically.  The cache is read as plain lists and ints only."""
    import pickle
    cache = particle_star + "_LocalSearch_Angle_" + str(int(local_range)) + "_data.pkl"
    entries = None
    if os.path.exists(cache):
        try:
            with open(cache, "rb") as fh:
                loaded = _PlainDataUnpickler(fh).load()
            if _local_cache_valid(loaded, particle_records, len(model_angles.rot)):
                entries = loaded
                print("Data loaded from file:", cache)
            else:
                print("Existing local-search cache is incompatible; rebuilding:", cache)
        except Exception as exc:  # unreadable or refused: rebuild
            print(f"Could not read local-search cache ({exc}); rebuilding: {cache}")
    if entries is None:
        entries = build_local_orientation_entries(particle_records, model_angles.rot, model_angles.tilt, local_range)
        tmp = cache + f".tmp.{os.getpid()}"
        try:
            with open(tmp, "wb") as fh:
                pickle.dump(entries, fh, protocol=pickle.HIGHEST_PROTOCOL)
            os.replace(tmp, cache)
            print("Data generated and saved to file:", cache)
        finally:
            if os.path.exists(tmp):
                try:
                    os.remove(tmp)
                except OSError:
                    pass
    return {int(e[0]): [int(x) for x in e[1:]] for e in entries}


def read_fsc_map(fsc_filename: str, original_box: int, new_box: int, ignore_fsc: bool) -> np.ndarray:
    """Centred [n, n] FSC weights (v606:471-491)."""
    return expand_radial_values(host.read_fsc_values(fsc_filename, original_box, ignore_fsc), new_box)


# ---------------------------------------------------------------------------
# Packed layout
# ---------------------------------------------------------------------------

class PackedFourierLayout:
    """Index maps of the packed rfft2 layout (v606:494-541)."""

    def __init__(self, size: int, crop_size: int, device):
        if size <= 0 or size % 2 != 0:
            raise ValueError("newboxsize must be a positive even integer")
        if crop_size <= 0 or crop_size % 2 != 0 or crop_size > size:
            raise ValueError("CCG crop size must be positive, even and <= newboxsize")
        self.size = int(size)
        self.rfft_size = self.size // 2 + 1
        self.crop_size = int(crop_size)
        self.device = torch.device(device)
        h = self.size // 2
        self.centered_row_to_packed = torch.cat((torch.arange(h, self.size), torch.arange(0, h))).to(self.device)
        self.centered_col_to_packed = torch.cat((torch.arange(h, self.size), torch.zeros(1, dtype=torch.long))).to(self.device)
        c = self.crop_size // 2
        self.unshifted_ccg_crop_indices = torch.cat((torch.arange(self.size - c, self.size),
                                                     torch.arange(0, c))).to(self.device)

    def centered_full_to_packed(self, centered: torch.Tensor) -> torch.Tensor:
        return centered.index_select(-2, self.centered_row_to_packed).index_select(-1, self.centered_col_to_packed)

    @staticmethod
    def pre_shift_for_correlation(packed: torch.Tensor) -> torch.Tensor:
        return torch.fft.fftshift(packed, dim=(-2, -1))

    @staticmethod
    def undo_pre_shift(pre_shifted: torch.Tensor) -> torch.Tensor:
        return torch.fft.ifftshift(pre_shifted, dim=(-2, -1))

    def centered_ccg_crop_from_unshifted(self, ccg: torch.Tensor) -> torch.Tensor:
        return ccg.index_select(-2, self.unshifted_ccg_crop_indices).index_select(-1, self.unshifted_ccg_crop_indices)


# ---------------------------------------------------------------------------
# CTF
# ---------------------------------------------------------------------------

class PackedCTFTorch:
    """Batched packed CTF on the device (v606:695-829), evaluated by gis_ctf_packed."""

    def __init__(self, size, pixel_size, cs, voltage, kk, device, amp_const=0.1, bfactor=0.0, scale=1.0,
                 phase_shift=0.0):
        if scale != 1.0:
            raise ValueError("scale != 1 is not used by the search path")
        self.size = int(size)
        self.device = torch.device(device)
        self.tables = host.CTFTables(size, pixel_size, cs, voltage, kk, amp_const, bfactor, phase_shift)
        self.k1, self.k2, self.k3, self.k5, self.kk = (self.tables.k1, self.tables.k2, self.tables.k3,
                                                       self.tables.k5, self.tables.kk)
        self._tab = torch.from_numpy(self.tables.table).to(self.device)

    def batch(self, records: Sequence[ParticleRecord], has_ctf: bool) -> torch.Tensor:
        B = len(records)
        if not has_ctf:
            return torch.ones((B, self.size, self.size // 2 + 1), dtype=torch.float32, device=self.device)
        dfo = torch.tensor([[r.defocus_u, r.defocus_v, r.defocus_angle] for r in records], dtype=torch.float32,
                           device=self.device)
        return kernels.ctf_packed(self._tab, dfo, self.size, self.tables.kconst)


class PackedCTFNumPy(PackedCTFTorch):
    """The reference's exact-layout backend interface (``--ctfBackend numpy``):
    same kernel, results returned as NumPy arrays."""

    def __init__(self, size, pixel_size, cs, voltage, kk, amp_const=0.1, bfactor=0.0, scale=1.0,
                 phase_shift=0.0, device=None):
        dev = torch.device("cuda", torch.cuda.current_device()) if device is None else device
        super().__init__(size, pixel_size, cs, voltage, kk, dev, amp_const, bfactor, scale, phase_shift)

    def one(self, defocus_u, defocus_v, defocus_angle) -> np.ndarray:
        return self.batch([ParticleRecord(0, "", 0.0, 0.0, 0.0, defocus_u, defocus_v, defocus_angle)], True)[0]

    def batch(self, records, has_ctf):  # type: ignore[override]
        return super().batch(records, has_ctf).cpu().numpy()


# ---------------------------------------------------------------------------
# Projector and particle preparation
# ---------------------------------------------------------------------------

class IntegratedProjector:
    """3-D transform + template bank on the HIP kernels (v606:1034-1155)."""

    def __init__(self, device, original_box, new_box, layout, real_space_mask, packed_fsc, projection_batch_size=8):
        self.device = torch.device(device)
        self.original_box = int(original_box)
        self.new_box = int(new_box)
        self.layout = layout
        self.real_space_mask = real_space_mask
        self.packed_fsc = packed_fsc
        self.batch_size = max(1, int(projection_batch_size))
        self.last_orientation_count = self.last_slice_count = 0

    def _volume(self, mrc_volume) -> torch.Tensor:
        if isinstance(mrc_volume, torch.Tensor):
            vol = mrc_volume
            if vol.dtype != torch.float32 or vol.device != self.device:
                raise ValueError("Model tensor must be float32 on the projector device")
        else:
            vol = torch.from_numpy(mrc.read(mrc_volume)).to(self.device)
        if vol.ndim != 3 or len(set(vol.shape)) != 1:
            raise ValueError(f"Model volume must be cubic; got shape {tuple(vol.shape)}")
        if vol.shape[0] != self.original_box:
            raise ValueError(f"Model volume box {vol.shape[0]} does not match --oriboxsize {self.original_box}")
        return vol.contiguous()

    def _selected(self, angles: ModelAngles, orientation_indices) -> Optional[List[int]]:
        if orientation_indices is None:
            return None
        sel = sorted(set(int(i) for i in orientation_indices))
        if any(i < 0 or i >= len(angles.rot) for i in sel):
            raise ValueError("Projection orientation index is outside the angle table")
        return sel

    def build_bank(self, mrc_volume, angles: ModelAngles, orientation_indices=None) -> torch.Tensor:
        """Unit-norm templates in the UNshifted packed layout (the engine's
        form), one row per orientation, or per sorted unique index of
        ``orientation_indices`` (the zip release's restricted projection,
        ZL v7:906-966; only those directions are sampled)."""
        sel = self._selected(angles, orientation_indices)
        mats = angles.rotation_matrices_t if sel is None else angles.rotation_matrices_t[sel]
        self.last_orientation_count = self.last_slice_count = int(len(mats))
        vol = self._volume(mrc_volume)
        if len(mats) == 0:
            return torch.empty((0, self.new_box, self.new_box // 2 + 1), dtype=torch.complex64, device=self.device)
        rts = torch.from_numpy(np.ascontiguousarray(mats, np.float32)).to(self.device)
        dft = kernels.volume_dft(vol)
        fsc = self.packed_fsc.to(torch.float32).contiguous()
        mask = None if self.real_space_mask is None else self.real_space_mask.to(torch.float32).contiguous()
        return kernels.build_templates(dft, self.new_box, rts, mask, fsc)

    def build_templates(self, mrc_volume, angles: ModelAngles, orientation_indices=None) -> torch.Tensor:
        """Pre-shifted bank [M, n, n/2+1] complex64, the v606 contract."""
        return self.layout.pre_shift_for_correlation(self.build_bank(mrc_volume, angles, orientation_indices))


class PsiGridBank:
    """Rotation grids of the searched psi angles (v606:1193-1256), stored as the
    six affine coefficients (c, -s, s, c, off_y, off_x) per angle that the
    preparation kernel samples with."""

    def __init__(self, size, device, unique_angles, build_batch_size=8):
        self.size = int(size)
        self.device = torch.device(device)
        self.angle_keys = [float(np.round(a, 6)) for a in unique_angles]
        self.angle_to_index = {a: i for i, a in enumerate(self.angle_keys)}
        self.params = torch.from_numpy(host.psi_grid_params(self.size, self.angle_keys)).to(self.device) \
            if self.angle_keys else torch.empty((0, 6), device=self.device)

    def indices_for_angles(self, effective_angles) -> torch.Tensor:
        idx = []
        for a in effective_angles:
            k = float(np.round(a, 6))
            try:
                idx.append(self.angle_to_index[k])
            except KeyError as exc:
                raise KeyError(f"Psi grid angle {k} was not precomputed") from exc
        return torch.tensor(idx, dtype=torch.long, device=self.device)

    def get(self, effective_angles) -> torch.Tensor:
        return self.params.index_select(0, self.indices_for_angles(effective_angles))


class ParticleMRCReader:
    """Reads particle images ``index@stack.mrcs`` into a device tensor (v606:1259-1318)."""

    def __init__(self, expected_size, device):
        self.expected_size = int(expected_size)
        self.device = torch.device(device)
        self._name = None
        self._data = None
        self.host_staging = torch.empty((self.expected_size, self.expected_size), dtype=torch.float32,
                                        pin_memory=True)

    def _open(self, filename):
        filename = os.path.realpath(filename) if os.path.islink(filename) else filename
        if filename != self._name:
            self._data = mrc.mmap(filename)
            self._name = filename

    def host_image(self, image_name: str) -> np.ndarray:
        serial, filename = image_name.split("@", 1)
        self._open(filename)
        d = self._data
        img = d if d.ndim < 3 else d[int(serial) - 1]
        if img.shape != (self.expected_size, self.expected_size):
            raise ValueError(f"Particle image shape {img.shape} does not match --oriboxsize "
                             f"{self.expected_size} in {image_name}")
        return np.asarray(img, dtype=np.float32)

    def read(self, image_name: str) -> torch.Tensor:
        self.host_staging.numpy()[...] = self.host_image(image_name)
        return self.host_staging.to(self.device)

    def read_many(self, image_names: Sequence[str]) -> torch.Tensor:
        """[B, N, N] device tensor of several images (one pinned upload)."""
        buf = torch.empty((len(image_names), self.expected_size, self.expected_size), dtype=torch.float32,
                          pin_memory=True)
        arr = buf.numpy()
        for i, name in enumerate(image_names):
            arr[i] = self.host_image(name)
        return buf.to(self.device, non_blocking=False)

    def close(self):
        self._data = None
        self._name = None


class ParticleFourierPreparer:
    """Particle psi stacks in the packed layout (v606:1321-1374) via gis_prepare_particles."""

    def __init__(self, original_box, new_box, device, layout, real_space_mask, psi_grid_bank):
        self.original_box = int(original_box)
        self.new_box = int(new_box)
        self.device = torch.device(device)
        self.layout = layout
        self.real_space_mask = real_space_mask
        self.psi_grid_bank = psi_grid_bank

    def prepare(self, image: torch.Tensor, effective_angles, packed_ctf: torch.Tensor) -> torch.Tensor:
        """Pre-shifted, unit-norm, conjugated spectra [P, n, n/2+1] (the v606 contract)."""
        rotp = self.psi_grid_bank.get(effective_angles).unsqueeze(0).contiguous()
        img = image.reshape(1, self.original_box, self.original_box).to(torch.float32).contiguous()
        ctf = packed_ctf.reshape(1, self.new_box, self.new_box // 2 + 1).to(torch.float32).contiguous()
        mask = None if self.real_space_mask is None else self.real_space_mask.to(torch.float32).contiguous()
        spec, inv = kernels.prepare_particles(img, self.new_box, rotp, mask, ctf, None, np.zeros(5, np.float32))
        return self.layout.pre_shift_for_correlation(spec[0] * inv[0, :, None, None])


# ---------------------------------------------------------------------------
# Correlation and records
# ---------------------------------------------------------------------------

class CorrelationWorkspace:
    """Kept for interface compatibility (v606:1382-1415); the kernel needs no chunk buffers."""

    def __init__(self, device):
        self.device = torch.device(device)


def correlate_templates_and_particle(templates_pre_shifted, particle_pre_shifted, layout, workspace=None,
                                     chunk_size_a=16, chunk_size_b=32):
    """(scores [A, B] float32, y [A, B] int64, x [A, B] int64) of v606:1418-1449."""
    A = int(templates_pre_shifted.shape[0])
    t = layout.undo_pre_shift(templates_pre_shifted).to(torch.complex64).contiguous()
    p = layout.undo_pre_shift(particle_pre_shifted).to(torch.complex64).contiguous().unsqueeze(0)
    dev = t.device
    off = torch.tensor([0, A], dtype=torch.int32, device=dev)
    idx = torch.arange(A, dtype=torch.int32, device=dev)
    c = kernels.correlate(t, p, None, off, idx, layout.crop_size, validate=False)
    return c.scores, c.peak_y.to(torch.int64), c.peak_x.to(torch.int64)


def make_result_records(particle, fake_psi, model_rot, model_tilt, model_indices, scores, y_positions,
                        x_positions, crop_size, trans_scale_factor, cc_threshold=3.0, accuracy=11.0):
    """Three records per particle (v606:1452-1516): cc, psi and rawCC lines.

    Mean and population std of the scores are taken in float64."""
    s = scores.detach().reshape(scores.shape[0], -1).cpu().numpy()
    if s.size == 0:
        return []
    ys = y_positions.detach().cpu().numpy().reshape(s.shape)
    xs = x_positions.detach().cpu().numpy().reshape(s.shape)
    flat = s.reshape(-1)
    k = int(np.argmax(flat))
    P = s.shape[1]
    li, pi = divmod(k, P)
    peak = float(s[li, pi])
    return _records(particle, fake_psi, model_rot, model_tilt, int(model_indices[li]), pi, peak,
                    int(ys[li, pi]), int(xs[li, pi]), float(flat.astype(np.float64).mean()),
                    float(flat.astype(np.float64).std()), crop_size, trans_scale_factor, cc_threshold, accuracy)


def _records(particle, fake_psi, model_rot, model_tilt, gi, pi, peak, py, px, mean, std, crop_size,
             trans_scale_factor, cc_threshold=3.0, accuracy=11.0):
    xshift = float(crop_size // 2 - px) / trans_scale_factor
    yshift = float(crop_size // 2 - py) / trans_scale_factor
    ad = calculate_angular_distance(model_rot[gi], model_tilt[gi], 0.0, particle.rot, particle.tilt, 0.0)
    z = (peak - mean) / std if std > 0 else 0.0
    return [
        ["cc: ", particle.line_index, float(z), 1 if z > cc_threshold else 0, 1 if ad < accuracy else 0,
         model_rot[gi], model_tilt[gi], ad, round(xshift), round(yshift), 0, 0],
        ["psi: ", particle.line_index, float(fake_psi[pi]), float((particle.psi + 360.0) % 360.0)],
        ["rawCC: ", particle.line_index, float(peak)],
    ]


# ---------------------------------------------------------------------------
# CLI and main search
# ---------------------------------------------------------------------------

def _nonnegative_mb(value):
    v = float(value)
    if not np.isfinite(v) or v < 0:
        raise argparse.ArgumentTypeError("must be a finite, non-negative MiB value")
    return v


def create_search_parser() -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(description="MI355X cryo-EM model/particle local search (GisAPR v606 interface)")
    p.add_argument("--i", type=str, required=True, help="Compatibility argument; model angles are read from --ang")
    p.add_argument("--p", type=str, required=True, help="Input particle STAR")
    p.add_argument("--FSC", type=str, default="", help="Input FSC file")
    p.add_argument("--o", type=str, required=True, help="Output search text file")
    p.add_argument("--kk", type=float, default=0.0)
    p.add_argument("--gpuid", type=int, default=0)
    p.add_argument("--oriboxsize", type=int, default=256)
    p.add_argument("--newboxsize", type=int, default=256)
    p.add_argument("--apix", type=float, default=1.42, help="Original particle pixel size")
    p.add_argument("--transRange", type=int, default=0, help="Compatibility argument; CCG peak determines translation")
    p.add_argument("--voltage", type=float, default=300.0)
    p.add_argument("--cs", type=float, default=2.7)
    p.add_argument("--maskRadius", type=int, default=110)
    p.add_argument("--maskEdge", type=int, default=6)
    p.add_argument("--ignoreFSC", action="store_true")
    p.add_argument("--discardMask", action="store_true")
    p.add_argument("--psiStep", type=float, default=15.0)
    p.add_argument("--start", type=int, default=-1)
    p.add_argument("--end", type=int, default=-1)
    p.add_argument("--doLocalSearch", action="store_true")
    p.add_argument("--localRange", type=float, default=20.0)
    p.add_argument("--doContinue", action="store_true")
    p.add_argument("--ang", type=str, required=True, help="RELION angle STAR used for projections and orientations")
    p.add_argument("--mrc", type=str, required=True, help="Input 3-D model MRC; retained on disk")
    p.add_argument("--ctfBackend", choices=("torch", "numpy"), default=os.environ.get("V606_CTF_BACKEND", "torch"))
    p.add_argument("--ctfBatchSize", type=int, default=256)
    p.add_argument("--projectionBatchSize", type=int, default=8)
    p.add_argument("--psiGridBuildBatch", type=int, default=8)
    p.add_argument("--ccgCropSize", type=int, default=32)
    p.add_argument("--chunkSizeA", type=int, default=16)
    p.add_argument("--chunkSizeB", type=int, default=32)
    p.add_argument("--outputBufferParticles", type=int, default=96)
    p.add_argument("--torchThreads", type=int, default=4)
    p.add_argument("--profileStages", action="store_true")
    # zip release: projection of the searched directions only, Top-K calibration banks
    p.add_argument("--projection_mode", choices=("needed", "full"), default="needed",
                   help="needed: project only the union of the searched directions; full: every direction")
    p.add_argument("--fixed_density_cache_mb", type=_nonnegative_mb, default=128.0)
    p.add_argument("--keep-orientations", dest="keep_orientations", type=orientation_bank.positive_int, default=3)
    g = p.add_mutually_exclusive_group()
    g.add_argument("--orientations-in", dest="orientations_in", metavar="JSON",
                   help="Score only the directions saved in this calibration bank.")
    g.add_argument("--orientations-out", dest="orientations_out", metavar="JSON",
                   help="Run calibration and export per-particle Top-K directions.")
    p.add_argument("--verbose", action="store_true")
    p.add_argument("--ctf_cache_mb", type=_nonnegative_mb, default=256.0)
    p.add_argument("--search_cache_mb", type=_nonnegative_mb, default=64.0)
    return p


def choose_device(gpuid) -> torch.device:
    """The HIP device for ``gpuid``; raises where there is none (no CPU path)."""
    if gpuid == "cpu" or not torch.cuda.is_available():
        raise RuntimeError("gisapr_amd runs on HIP devices only; no GPU is visible")
    torch.cuda.set_device(int(gpuid))
    return torch.device(f"cuda:{int(gpuid)}")


def sync_if_cuda(device: torch.device) -> None:
    if device.type == "cuda":
        torch.cuda.synchronize(device)


def chunks(sequence, size) -> Iterable:
    size = max(1, int(size))
    for start in range(0, len(sequence), size):
        yield sequence[start:start + size]


def _session_get(session_cache, key, factory):
    """Parsed STAR tables and angle lists kept across requests: the worker's
    SearchCache (its ``metadata.get_or_build``, Z3 gisapr_cache.py:110-117,
    120-123), a plain dict, or nothing (always parse)."""
    if session_cache is None:
        return factory()
    store = getattr(session_cache, "metadata", None)
    if store is not None and hasattr(store, "get_or_build"):
        return store.get_or_build(key, factory)
    if key not in session_cache:
        session_cache[key] = factory()
    return session_cache[key]


def run_search(args: argparse.Namespace, *, model_volume=None, session_cache=None, orientation_rows=None,
               calibrate=False) -> Dict[str, object]:
    """v606 run_search: writes ``args.o`` (three records per particle) and returns
    timing metrics (search_s, projection_s, setup_s, execution_s, ctf_hits).

    The zip release's additions (ZL v7:1483-1790): ``orientation_rows`` (a
    calibration bank's rows, or --orientations-in) restricts each particle to
    its bank directions; ``calibrate`` (or --orientations-out, which also
    writes the bank) ranks each particle's directions and returns them as
    ``orientation_rows``; with --projection_mode needed only the union of the
    searched directions is projected (``template_directions``)."""
    total_start = time.time()
    verbose = bool(getattr(args, "verbose", False))

    def log(*v, **kw):
        if verbose:
            print(*v, **kw)

    def meta(key, factory):
        return _session_get(session_cache, key, factory)

    device = choose_device(args.gpuid)
    N, n = int(args.oriboxsize), int(args.newboxsize)
    if N % 2 or n % 2:
        raise ValueError("oriboxsize and newboxsize must be even")
    if n > N:
        raise ValueError("newboxsize cannot exceed oriboxsize")
    if args.psiStep <= 0:
        raise ValueError("psiStep must be positive")
    layout = PackedFourierLayout(n, int(args.ccgCropSize), device)

    def fkey(path):
        st = os.stat(path)
        return (os.path.realpath(path), st.st_size, st.st_mtime_ns)

    angles = meta(("angles", fkey(args.ang)), lambda: read_model_angles(args.ang))
    table = meta(("particles", fkey(args.p)), lambda: read_particle_table(args.p))
    selected = select_particle_records(table, args.start, args.end)
    log("Model orientation count:", len(angles.rot), "Selected particle count:", len(selected))
    keep = int(getattr(args, "keep_orientations", 3))
    bank_in, bank_out = getattr(args, "orientations_in", None), getattr(args, "orientations_out", None)
    context = None
    if bank_in or bank_out:
        context = orientation_bank.settings(args.p, args.ang, keep, args.doLocalSearch, args.localRange,
                                            args.psiStep)
    if bank_in:
        saved = json.loads(Path(bank_in).read_text(encoding="utf-8"))
        orientation_bank.check_bank(saved, context)
        orientation_rows = saved["rows"]
    calibrate = bool(calibrate) or bool(bank_out)
    if orientation_rows is not None and calibrate:
        raise ValueError("Cannot calibrate from already restricted directions")
    fsc_values = host.read_fsc_values(args.FSC, N, bool(args.ignoreFSC))
    engine = SearchEngine(box=N, new_box=n, apix=float(args.apix), crop=int(args.ccgCropSize),
                          psi_step=float(args.psiStep), local_range=float(args.localRange),
                          do_local_search=bool(args.doLocalSearch), kk=float(args.kk), voltage=float(args.voltage),
                          cs=float(args.cs), mask_radius=int(args.maskRadius), mask_edge=int(args.maskEdge),
                          discard_mask=bool(args.discardMask), fsc_values=fsc_values, dir_rot=angles.rot,
                          dir_tilt=angles.tilt, device=device)

    # searched directions per particle: the local map, restricted to a bank's rows when given
    local_all = engine.local_directions([r.rot for r in selected], [r.tilt for r in selected])
    if orientation_rows is not None:
        local_all = [orientation_bank.restricted_indices(orientation_rows.get(str(rec.line_index)), rec, allowed,
                                                         angles, keep)
                     for rec, allowed in zip(selected, local_all)]
    needed = None
    if getattr(args, "projection_mode", "needed") == "needed":
        needed = sorted({i for l in local_all for i in l})

    projection_start = time.time()
    projector = IntegratedProjector(device, N, n, layout, engine.mask, engine.fsc_packed)
    volume_src = args.mrc if model_volume is None else model_volume
    if needed is None:
        bank = projector.build_bank(volume_src, angles)
    else:  # rows of unsearched directions stay unset: no particle reads them
        bank = torch.empty((len(angles.rot), n, n // 2 + 1), dtype=torch.complex64, device=device)
        if needed:
            rows = projector.build_bank(volume_src, angles, needed)
            bank.index_copy_(0, torch.as_tensor(needed, dtype=torch.long, device=device), rows)
        else:
            projector.last_orientation_count = projector.last_slice_count = 0
    sync_if_cuda(device)
    projection_seconds = time.time() - projection_start
    ranked: Dict[str, dict] = {}

    search_start = time.time()
    reader = ParticleMRCReader(N, device)
    output_mode = "a" if args.doContinue else "w"
    try:
        with open(args.o, output_mode, encoding="utf-8") as out:
            group_size = max(1, int(args.ctfBatchSize)) * 8
            for g0 in range(0, len(selected), group_size):
                group = selected[g0:g0 + group_size]
                local = local_all[g0:g0 + group_size]
                images = reader.read_many([r.image_name for r in group])
                dfo = None
                if table.has_ctf:
                    dfo = np.asarray([[r.defocus_u, r.defocus_v, r.defocus_angle] for r in group], np.float32)
                ps = engine.add_particles(images, [r.rot for r in group], [r.tilt for r in group],
                                          [r.psi for r in group], defocus=dfo,
                                          line_index=[r.line_index for r in group], local=local)
                res = engine.search(bank, ps)
                recs = {r["line"]: r for r in engine.records(res, ps)}
                if calibrate:
                    scores = res.corr.scores.reshape(-1, ps.P).cpu().numpy()
                    off = ps.dir_off.cpu().numpy()
                    for b, rec in enumerate(group):
                        ranked[str(rec.line_index)] = orientation_bank.ranked_row(
                            rec, local[b], angles, scores[off[b]:off[b + 1]], keep)
                lines: List[str] = []
                for b, rec in enumerate(group):
                    r = recs.get(rec.line_index)
                    if r is None:  # no model orientation within the local range
                        continue
                    for item in _records(rec, ps.fake_psi[b], angles.rot, angles.tilt, r["direction"],
                                         r["best_psi"], r["rawcc"], r["peak_y"], r["peak_x"], r["mean"], r["std"],
                                         engine.crop, engine.trans_scale):
                        lines.append(str(item) + "\n")
                out.writelines(lines)
                out.flush()
    finally:
        reader.close()
    sync_if_cuda(device)
    search_seconds = time.time() - search_start
    total = time.time() - total_start
    log("Total search time =", round(search_seconds, 4), "seconds")
    if bank_out:
        if any(row.get("invalid_rawcc") for row in ranked.values()):
            raise ValueError("NaN rawCC in calibration: no usable orientation bank can be exported")
        orientation_bank.atomic_json(bank_out, orientation_bank.pack_bank(context, ranked,
                                                                          mrc=str(Path(args.mrc).resolve())))
    return dict(orientation_rows=ranked if calibrate else None, search_s=search_seconds,
                projection_s=projection_seconds, setup_s=max(0.0, total - search_seconds - projection_seconds),
                execution_s=total, ctf_hits=0, template_directions=projector.last_orientation_count,
                sampled_directions=projector.last_slice_count)


def main() -> None:
    if "--worker-connect" in sys.argv[1:]:
        from .worker import run_worker
        run_worker(create_search_parser, run_search, choose_device)
        return
    args = create_search_parser().parse_args()
    with torch.inference_mode():
        run_search(args)


if __name__ == "__main__":
    main()
