"""Batched multi-candidate evaluation: one optimizer step on the device.

A PSO swarm, a simplex or a grid step evaluates many rigid-body candidates of
one moving chain against the same particle stack.  ``CandidateBatch`` keeps
everything of such a step resident in HBM and runs it as a fixed sequence of
launches that can be captured once as a HIP graph and replayed per step:

  all slots at once:     the moving chain moved by each slot's pose and
                         8.3-rounded (gis_move_chain)
  per candidate slot c:  fixed-body density (copy)  ->  splat of the moving
                         chain  ->  3-D transform  ->  768-direction template
                         bank  ->  (small-shift path) in-disk coefficients
  all slots at once:     one search launch over the virtual particles
                         (slot c, particle b in the slot's subset), reading the
                         slot's bank rows c*M + direction and the shared
                         particle spectra through a particle map
                         one segmented float64 score sum per slot.

The only host work per step is each slot's rotation matrix and translation
(12 doubles; the rigid-body move and 8.3 rounding of ZL:gisapr_batch_density.py:
41-49 then run on the device) and one host-to-device copy of them.

``RigidDensityCache`` and ``BatchSearchEngine`` keep the call surface of
ZL:gisapr_batch_density.py:18-117 and ZL:gisapr_batch_search.py:142-364.
``Autotuner`` times the search kernel's tile variants on the real operands
once per shape and keeps the fastest (in memory and, optionally, in a JSON
file).
"""
from __future__ import annotations

import json
import os
import time
from dataclasses import dataclass
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np
import torch

from . import kernels
from .density import atoms_to_atomic_numbers_and_masses, density_from_coords, splat_parameters, yflip_vectorized
from .engine import ParticleSet, SearchEngine
from .pdb_text import quantize_8_3, read_pdb_cached, rotvec_matrix, transform_coordinates


# ---------------------------------------------------------------------------
# Autotuner
# ---------------------------------------------------------------------------

class Autotuner:
    """Chooses the fastest tile variant of a launch on its real operands.

    ``choose(key, variants, run)`` times ``run(variant)`` for each variant
    (one untimed call, then ``reps`` timed calls with HIP events on the
    current stream) and returns the fastest; the choice is remembered per
    ``key`` and written to ``path`` (JSON) when one is given.
    """

    def __init__(self, path: Optional[str] = None, reps: int = 2):
        self.path = path
        self.reps = int(reps)
        self.table: Dict[str, dict] = {}
        self.timings: Dict[str, List[tuple]] = {}
        if path and os.path.exists(path):
            with open(path, "r", encoding="utf-8") as fh:
                self.table = json.load(fh)

    @staticmethod
    def key(*parts) -> str:
        return "/".join(str(p) for p in parts)

    def choose(self, key: str, variants: Sequence[dict], run: Callable[[dict], object]) -> dict:
        if key in self.table:
            return dict(self.table[key])
        results = []
        for v in variants:
            run(v)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(self.reps):
                run(v)
            e1.record()
            torch.cuda.synchronize()
            results.append((e0.elapsed_time(e1) / self.reps, v))
        best = min(results, key=lambda r: r[0])[1]
        self.table[key] = dict(best)
        self.timings[key] = [(ms, dict(v)) for ms, v in results]
        if self.path:
            tmp = self.path + ".tmp"
            with open(tmp, "w", encoding="utf-8") as fh:
                json.dump(self.table, fh, indent=1, sort_keys=True)
            os.replace(tmp, self.path)
        return dict(best)


MIXED_VARIANTS = [dict(pt=5, wpb=4), dict(pt=5, wpb=2), dict(pt=3, wpb=4), dict(pt=1, wpb=4)]
SPLIT_VARIANTS = [dict(dt=2, pt=5), dict(dt=1, pt=5), dict(dt=4, pt=5), dict(dt=2, pt=3)]
SPLIT16_VARIANTS = [dict(dt=1, pt=5), dict(dt=2, pt=5), dict(dt=1, pt=3), dict(dt=2, pt=2)]
PATHS = ("fp32", "split", "split16", "mixed")
FP32_VARIANTS = [dict(dt=1, pt=5), dict(dt=2, pt=5), dict(dt=1, pt=3), dict(dt=2, pt=3)]


def _variants_for(path: str, P: int) -> List[dict]:
    table = {"mixed": MIXED_VARIANTS, "split": SPLIT_VARIANTS, "split16": SPLIT16_VARIANTS}.get(path, FP32_VARIANTS)
    out = []
    for v in table:
        v = dict(v, pt=min(v["pt"], P))
        if v not in out:
            out.append(v)
    return out


# ---------------------------------------------------------------------------
# Search plan over candidate slots
# ---------------------------------------------------------------------------

@dataclass
class SlotPlan:
    """Virtual particles (slot c, particle b) of one batched search launch."""
    slots: int
    subsets: List[np.ndarray]        # particle indices per slot (ascending)
    dir_off: torch.Tensor            # [V+1] int32
    dir_idx: torch.Tensor            # [sum] int32: c * M + direction
    pmap: torch.Tensor               # [V] int32: particle of each virtual particle
    slot_off: torch.Tensor           # [slots+1] int32 virtual-particle offsets
    slot_off_host: np.ndarray
    row_off_host: np.ndarray         # [slots+1] dir_idx offsets at slot boundaries
    n_pairs: int                     # (direction, psi) pairs searched per step

    row_particle: np.ndarray = None  # [rows] particle of each dir_idx row
    _orders: dict = None

    def prefix(self, k: int):
        """Operands of a search over the first k slots only (views, no copies)."""
        v = int(self.slot_off_host[k])
        r = int(self.row_off_host[k])
        return self.dir_off[:v + 1], self.dir_idx[:r], self.pmap[:v], self.slot_off[:k + 1]

    particle_rank: np.ndarray = None  # [particles] processing rank of each particle
    pmap_host: np.ndarray = None      # [V] particle of each virtual particle

    def particle_major(self, k: int) -> torch.Tensor:
        """Processing order of the first k slots' rows: each particle's rows of all
        slots adjacent (they share the particle's coefficients in cache), particles
        in ``particle_rank`` order (neighbours in orientation share templates)."""
        if self._orders is None:
            self._orders = {}
        if k not in self._orders:
            r = int(self.row_off_host[k])
            key = self.particle_rank[self.row_particle[:r]]
            order = np.argsort(key, kind="stable").astype(np.int32)
            self._orders[k] = torch.from_numpy(order).to(self.dir_idx.device)
        return self._orders[k]

    def virtual_major(self, k: int) -> torch.Tensor:
        """Processing order of the first k slots' virtual particles: the slots'
        copies of one particle adjacent (they share its spectra in cache),
        particles in ``particle_rank`` order."""
        if self._orders is None:
            self._orders = {}
        key = ("virtual", k)
        if key not in self._orders:
            v = int(self.slot_off_host[k])
            pm = self.pmap_host[:v]
            order = np.argsort(self.particle_rank[pm], kind="stable").astype(np.int32)
            self._orders[key] = torch.from_numpy(order).to(self.dir_idx.device)
        return self._orders[key]

    @property
    def virtual(self) -> int:
        return int(self.slot_off_host[-1])


def orientation_rank(dir_off: np.ndarray, dir_idx: np.ndarray) -> np.ndarray:
    """Rank of each particle when sorted by its first local direction (particles
    whose priors are close share most of their local directions)."""
    off = np.asarray(dir_off, np.int64)
    idx = np.asarray(dir_idx, np.int64)
    first = np.where(off[1:] > off[:-1], idx[np.minimum(off[:-1], max(len(idx) - 1, 0))], -1)
    order = np.argsort(first, kind="stable")
    rank = np.empty(len(first), np.int64)
    rank[order] = np.arange(len(first))
    return rank


def build_slot_plan(ps: ParticleSet, M: int, slots: int, subsets: Optional[Sequence] = None,
                    device=None) -> SlotPlan:
    """Plan the batched search of ``slots`` candidates; slot c searches the
    particles ``subsets[c]`` (all particles when None)."""
    off = ps.dir_off.cpu().numpy().astype(np.int64)
    idx = ps.dir_idx.cpu().numpy().astype(np.int64)
    B = ps.B
    if subsets is None:
        subsets = [np.arange(B)] * slots
    if len(subsets) != slots:
        raise ValueError("one particle subset per slot is required")
    subs, v_off, v_idx, v_map, s_off = [], [0], [], [], [0]
    for c, sub in enumerate(subsets):
        sub = np.unique(np.asarray(sub, np.int64))
        if len(sub) and (sub[0] < 0 or sub[-1] >= B):
            raise ValueError("particle subset index out of range")
        subs.append(sub)
        counts = off[sub + 1] - off[sub]
        starts = off[sub]
        rows = np.repeat(starts, counts) + (np.arange(counts.sum()) - np.repeat(np.cumsum(counts) - counts, counts))
        v_idx.append(idx[rows] + c * M)
        v_off.append(counts)
        v_map.append(sub)
        s_off.append(len(sub))
    d_off = np.concatenate([[0], np.cumsum(np.concatenate(v_off[1:]) if len(v_off) > 1 else [])]).astype(np.int32)
    row_particle = np.repeat(np.concatenate(v_map) if v_map else np.zeros(0, np.int64),
                             np.concatenate(v_off[1:]) if len(v_off) > 1 else np.zeros(0, np.int64))
    d_idx = np.concatenate(v_idx).astype(np.int32) if v_idx else np.zeros(0, np.int32)
    pm = np.concatenate(v_map).astype(np.int32) if v_map else np.zeros(0, np.int32)
    so = np.cumsum(s_off).astype(np.int32)
    dev = ps.dir_off.device if device is None else device
    return SlotPlan(slots=slots, subsets=subs, dir_off=torch.from_numpy(d_off).to(dev),
                    dir_idx=torch.from_numpy(d_idx).to(dev), pmap=torch.from_numpy(pm).to(dev),
                    slot_off=torch.from_numpy(so).to(dev), slot_off_host=so, row_off_host=d_off[so].copy(),
                    n_pairs=int(len(d_idx)) * ps.P, row_particle=row_particle,
                    particle_rank=orientation_rank(off, idx), pmap_host=pm)


# ---------------------------------------------------------------------------
# Fixed body + moving chain
# ---------------------------------------------------------------------------

@dataclass
class RigidBody:
    """A model split into a fixed body and one moving chain (first-model rows)."""
    fixed_coords: np.ndarray         # [F, 3] float32
    fixed_numbers: np.ndarray        # [F]
    moving_coords: np.ndarray        # [A, 3] float32, first-model coordinates of the moving chain
    moving_numbers: np.ndarray       # [A]

    @classmethod
    def from_arrays(cls, coords, numbers, chains, moving_chain):
        chains = np.asarray(chains)
        sel = chains == str(moving_chain)
        if not np.any(sel):
            raise ValueError(f"chain {moving_chain!r} has no atoms")
        coords = np.asarray(coords, np.float32)
        numbers = np.asarray(numbers, np.float32)
        return cls(coords[~sel], numbers[~sel], coords[sel], numbers[sel])

    @classmethod
    def from_pdb(cls, source_pdb, chain_id, include_hetatm=False):
        """The split RigidDensityCache makes of a PDB file: the moving chain is
        the first model's atoms of ``chain_id``; every other atom of every
        model (with HETATM records when ``include_hetatm``) is fixed."""
        first = read_pdb_cached(source_pdb)
        dens = read_pdb_cached(source_pdb, include_hetatm=include_hetatm, first_model_only=False)
        sel = first.chains == str(chain_id)
        if not sel.any():
            raise ValueError(f"Chain {chain_id!r} has no atoms in the first PDB model")
        moving_lines = first.line_numbers[sel]
        moving = np.isin(dens.line_numbers, moving_lines)
        numbers, _ = atoms_to_atomic_numbers_and_masses(dens.elements)
        numbers = np.asarray(numbers, np.float32)
        coords = np.asarray(dens.coords, np.float32)
        return cls(coords[~moving], numbers[~moving], coords[moving], numbers[moving])

    def moved(self, pose) -> np.ndarray:
        """Moving-chain coordinates after ``pose`` (rotvec deg, translation A) about
        the chain's mean, 8.3-rounded (pdb_text.py:123-142)."""
        return quantize_8_3(transform_coordinates(self.moving_coords, pose))

    @property
    def centre(self) -> np.ndarray:
        """The pivot ``moved`` uses: the float32 mean of the chain's coordinates."""
        return np.asarray(self.moving_coords).mean(axis=0)

    def pose_rt(self, pose) -> np.ndarray:
        """float64 [12] operand of ``kernels.move_chain`` for ``pose``: the rotation
        matrix ``moved`` applies (row-major), then the translation.  Poses whose
        moved chain may leave the PDB 8.3 field are checked on the host first
        (``moved`` raises for them, as writing the coordinates would)."""
        p = np.asarray(pose, np.float64)
        if p.shape != (6,) or not np.isfinite(p).all():
            raise ValueError("pose must be six finite numbers: rotation vector (deg) and translation (A)")
        if not hasattr(self, "_radius"):
            d = np.asarray(self.moving_coords, np.float64) - self.centre
            self._radius = float(np.sqrt((d * d).sum(axis=1).max())) if len(d) else 0.0
        base = self.centre.astype(np.float64) + p[3:]
        # a rotation keeps every |x - centre| <= radius; 0.01 A covers the rounding
        if np.any(base + self._radius + 0.01 >= 9999.9995) or np.any(base - self._radius - 0.01 <= -999.9995):
            self.moved(p)  # raises when a coordinate really is outside the field
        return np.concatenate([rotvec_matrix(p[:3]).reshape(9), p[3:]])


# ---------------------------------------------------------------------------
# The batched step
# ---------------------------------------------------------------------------

class CandidateBatch:
    """``slots`` candidate poses of one moving chain, searched against one
    resident particle set, on the FP32 ("fp32"), full-plane split-bf16
    ("split"), full-plane fp16 with split-bf16 rescoring of each particle's
    best pair and near ties ("split16") or split-bf16 small-shift ("mixed")
    path.

    ``evaluate(poses)`` runs one step eagerly; ``capture()`` records the
    device part as a HIP graph and later ``evaluate`` calls replay it.
    Returns the float64 score sum_b exp(rawCC_b / 1024) per candidate.

    The banks of ``template_group`` candidates are built by one template
    launch over their spectra (each spectrum stays resident until then: one
    3-D half spectrum per grouped candidate).  The default groups just
    enough candidates to give every resident template workgroup a direction
    (one candidate when the direction table alone fills the device).
    """

    def __init__(self, engine: SearchEngine, ps: ParticleSet, body: RigidBody, res: float, slots: int, *,
                 path: str = "mixed", subsets=None, autotuner: Optional[Autotuner] = None,
                 template_group: Optional[int] = None):
        if path not in PATHS:
            raise ValueError(f"path must be one of {PATHS}")
        self.engine, self.ps, self.body, self.path = engine, ps, body, path
        self.slots = int(slots)
        if template_group is None:
            resident = int(kernels.lib().gis_build_templates_resident(engine.N, 1))
            template_group = min(8, max(1, -(-resident // max(1, len(engine.dir_rot)))))
        if int(template_group) < 1:
            raise ValueError("template_group must be >= 1")
        self.group = min(int(template_group), self.slots)
        dev = engine.device
        self.device = dev
        N, n = engine.N, engine.n
        M = len(engine.dir_rot)
        self.M = M
        self.res = float(res)
        self.width, self.rp, self.kn = splat_parameters(engine.apix, res)
        if path == "mixed":
            if getattr(engine, "small", None) is None:
                engine.enable_small_shift()
            if ps.coeffs is None:
                engine.prepare_small_shift(ps)
        elif ps.spectra_frag is None:
            engine.prepare(ps)
        self.plan = build_slot_plan(ps, M, self.slots, subsets, dev)
        # resident buffers
        self.fixed = density_from_coords(body.fixed_coords, body.fixed_numbers, N, engine.apix, res, device=dev)
        self.moving_numbers = torch.as_tensor(body.moving_numbers, dtype=torch.float32).to(dev).contiguous()
        A = len(body.moving_numbers)
        # per slot the pose (rotation matrix and translation); the step moves the chain on the device
        self.rt_host = torch.empty((self.slots, 12), dtype=torch.float64).pin_memory()
        self.rt_dev = torch.zeros((self.slots, 12), dtype=torch.float64, device=dev)
        self.rt_dev[:, 0] = self.rt_dev[:, 4] = self.rt_dev[:, 8] = 1.0
        self.moving_dev = torch.as_tensor(np.ascontiguousarray(body.moving_coords, np.float32)).to(dev)
        self.centre = body.centre
        self.coords_dev = torch.zeros((self.slots, A, 3), dtype=torch.float32, device=dev)
        self.volume = torch.empty((N, N, N), dtype=torch.float32, device=dev)
        self.dft = torch.empty((self.group, int(kernels.lib().gis_volume_dft_bytes(N)) // 8), dtype=torch.complex64,
                               device=dev)
        self.dft_ws = torch.empty((int(kernels.lib().gis_volume_dft_workspace(N)),), dtype=torch.uint8, device=dev)
        self.bank16 = None
        if path == "mixed":
            # template coefficients written straight from the template kernel (no bank)
            self.bank = None
            self.tk = torch.empty((self.slots * M, engine.small.Kpad), dtype=torch.complex64, device=dev)
            self.norms = torch.empty((self.slots * M,), dtype=torch.float32, device=dev)
        else:  # the searches' operand layout, written by the template kernel
            self.bank = kernels.FragmentPlanes(torch.empty((self.slots * M, kernels.fragment_plane(n)),
                                                           dtype=torch.complex64, device=dev), n)
            engine.search_spectra(ps)
            if path == "split16":  # the bank as fp16 pairs for the fp16 pass, packed after each template launch
                engine.search_spectra16(ps)
                self.bank16 = kernels.HalfFragmentPlanes(torch.empty((self.slots * M, kernels.fragment_plane(n)),
                                                                     dtype=torch.complex32, device=dev), n)
        self.tinv = torch.empty((self.slots * M,), dtype=torch.float32, device=dev)  # 1/norm per bank row
        self.tpl_ws = torch.empty((max(int(kernels.lib().gis_build_templates_workspace(N, self.group * M)), 1),),
                                  dtype=torch.uint8, device=dev)
        self._copied = torch.cuda.Event()
        self.autotuner = autotuner
        self.tile: dict = {}
        self.graph = None
        self._graph_tables = None
        self.static_out = None

    # -- device part ---------------------------------------------------------
    def _spectrum(self, c: int, g: int):
        """Slot c's density (resident fixed body + splatted moving chain) and
        its 3-D half spectrum into group position g."""
        eng = self.engine
        self.volume.copy_(self.fixed)
        kernels.splat(self.coords_dev[c], self.moving_numbers, eng.N, self.width, eng.apix, self.rp, self.kn,
                      self.volume)
        kernels.volume_dft(self.volume, out=self.dft[g], workspace=self.dft_ws)

    def _templates(self, c0: int, c1: int):
        """Banks (or in-disk coefficients) of slots [c0, c1), one template launch."""
        eng = self.engine
        for c in range(c0, c1):
            self._spectrum(c, c - c0)
        dft = self.dft[:c1 - c0]
        rows = slice(c0 * self.M, c1 * self.M)
        if self.path == "mixed":
            kernels.build_template_coefficients(dft, eng.n, eng.rts, eng.mask, eng.fsc_packed, eng.small_kslot,
                                                eng.small.K, eng.small.Kpad, out=self.tk[rows],
                                                norms=self.norms[rows], workspace=self.tpl_ws)
            return
        _, norms = kernels.build_templates(dft, eng.n, eng.rts, eng.mask, eng.fsc_packed, out=self.bank[rows],
                                           normalize=False, workspace=self.tpl_ws)
        torch.reciprocal(norms, out=self.tinv[rows])
        if self.bank16 is not None:
            kernels.pack_fragments_f16(self.bank[rows], self.tinv[rows], out=self.bank16.data[rows])

    def _all_templates(self, k: int):
        kernels.move_chain(self.moving_dev, self.centre, self.rt_dev[:k], out=self.coords_dev[:k])
        for c0 in range(0, k, self.group):
            self._templates(c0, min(c0 + self.group, k))

    def _search(self, tile: dict, k: Optional[int] = None):
        eng, ps = self.engine, self.ps
        dir_off, dir_idx, pmap, _ = self.plan.prefix(self.slots if k is None else k)
        if self.path == "mixed":
            return kernels.correlate_mixed(self.tk, ps.coeffs, ps.P, dir_off, dir_idx, eng.small_bhi,
                                           eng.small_blo, eng.small.half, validate=False, pmap=pmap,
                                           pt=tile.get("pt", 0), wpb=tile.get("wpb", 0), plast=eng.small_plast,
                                           rows=self.plan.particle_major(self.slots if k is None else k),
                                           operand=eng.small.operand)
        if self.path == "split":
            return kernels.correlate_split(self.bank, eng.search_spectra(ps), ps.inv_norm, dir_off, dir_idx, eng.crop,
                                           validate=False, pmap=pmap, tinv=self.tinv,
                                           order=self.plan.virtual_major(self.slots if k is None else k),
                                           dt=tile.get("dt", 0), pt=tile.get("pt", 0))
        if self.path == "split16":
            return kernels.correlate_split_f16(self.bank, eng.search_spectra16(ps), dir_off, dir_idx, eng.crop,
                                               validate=False, pmap=pmap, tinv=self.tinv,
                                               order=self.plan.virtual_major(self.slots if k is None else k),
                                               dt=tile.get("dt", 0), pt=tile.get("pt", 0),
                                               refine=eng.search_spectra(ps), inv_norm=ps.inv_norm,
                                               bank16=self.bank16)
        return kernels.correlate(self.bank, eng.search_spectra(ps), ps.inv_norm, dir_off, dir_idx, eng.crop,
                                 validate=False, pmap=pmap, tinv=self.tinv,
                                 order=self.plan.virtual_major(self.slots if k is None else k),
                                 dt=tile.get("dt", 0), pt=tile.get("pt", 0))

    def _device_step(self, k: Optional[int] = None):
        k = self.slots if k is None else k
        self._all_templates(k)
        corr = self._search(self.tile, k)
        scores = kernels.score_sum_segments(corr.peak, corr.best, self.plan.prefix(k)[3])
        return scores, corr

    # -- host part -----------------------------------------------------------
    def _stage(self, poses):
        poses = np.asarray(poses, np.float64).reshape(-1, 6)
        if len(poses) > self.slots:
            raise ValueError(f"{len(poses)} poses for {self.slots} slots")
        k = len(poses)
        fill = self.slots if self.graph is not None else k  # a graph runs every slot: repeat the last pose
        self._copied.synchronize()  # the previous step's copy out of the pinned buffer has run
        for c in range(fill):
            self.rt_host[c].copy_(torch.from_numpy(self.body.pose_rt(poses[min(c, k - 1)])))
        self.rt_dev[:fill].copy_(self.rt_host[:fill], non_blocking=True)
        self._copied.record()
        return k

    def tune(self, poses):
        """Pick the search kernel's tile variant for this plan on the banks of ``poses``."""
        if self.autotuner is None:
            return self.tile
        self._stage(poses)
        self._all_templates(self.slots)
        ps, pl = self.ps, self.plan
        if self.path == "mixed":
            key = Autotuner.key("mixed" if self.engine.small.operand == "bf16" else "mixed_fp16", self.engine.n,
                                ps.P, self.engine.small.Kpad, _bucket(int(pl.dir_idx.numel())))
        else:
            key = Autotuner.key(self.path, self.engine.n, ps.P, _bucket(int(pl.dir_idx.numel())))
        self.tile = self.autotuner.choose(key, _variants_for(self.path, ps.P), self._search)
        return self.tile

    def capture(self, poses=None):
        """Record the device part of a step as a HIP graph (after one eager warm-up step)."""
        if poses is not None:
            self._stage(poses)
        torch.cuda.synchronize()
        self._device_step()  # warm-up: library set-up calls happen outside the capture
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.static_out = self._device_step()
        self.graph = g
        # the graph holds the device addresses of every operand it reads: keep
        # those tensors alive and refuse replays once any of them is replaced
        self._graph_tables = self._graph_operands()
        return g

    def _graph_operands(self):
        """The engine / particle-set tensors a captured step reads, by identity."""
        eng, ps = self.engine, self.ps
        if self.path == "mixed":
            return (ps.coeffs, eng.small_bhi, eng.small_blo, eng.small_plast, eng.small_kslot)
        if self.path == "split16":
            return (ps.spectra_frag.data, ps.inv_norm, ps.spectra16.data)
        return (ps.spectra_frag.data, ps.inv_norm)

    def evaluate(self, poses, sync: bool = True):
        """Score ``poses`` ([k, 6], k <= slots); returns float64 [k] (or the device
        tensor and the CorrelationResult when sync is False).  Eager steps run
        the first k slots only; a captured graph always runs every slot."""
        k = self._stage(poses)
        if self.graph is not None:
            now = self._graph_operands()
            if len(now) != len(self._graph_tables) or any(a is not b for a, b in zip(now, self._graph_tables)):
                raise RuntimeError("operands of the captured step were replaced since capture(); capture again")
            self.graph.replay()
            scores, corr = self.static_out
        else:
            scores, corr = self._device_step(k)
        if not sync:
            return scores, corr
        return scores.cpu().numpy()[:k].copy()

    def slot_result(self, corr, c: int):
        """Per-particle view (best, peak, mean, std) of slot c's virtual particles."""
        a, b = int(self.plan.slot_off_host[c]), int(self.plan.slot_off_host[c + 1])
        return dict(particles=self.plan.subsets[c], best=corr.best[a:b], peak=corr.peak[a:b], mean=corr.mean[a:b],
                    std=corr.std[a:b])


def _bucket(x: int) -> int:
    return 1 << max(0, int(x - 1).bit_length())


# ---------------------------------------------------------------------------
# Reference-shaped front ends
# ---------------------------------------------------------------------------

class RigidDensityCache:
    """Fixed-body density reuse (ZL:gisapr_batch_density.py:18-117).

    The fixed atoms (all density rows outside the moving chain's first-model
    lines) are splatted once and kept on the device; each ``generate`` clones
    them and splats only the moving chain after the rigid-body transform and
    8.3 rounding.  Atomic accumulation order differs from a single splat of
    all atoms: equal up to float32 rounding, not bitwise."""

    def __init__(self):
        self.key = None
        self.fixed = None
        self.last_metrics = {}

    def clear(self):
        self.key = self.fixed = None
        self.moving_coords = self.moving_numbers = self.transform_rows = None

    @staticmethod
    def source_key(kwargs, max_mb, mode, backend):
        source = kwargs.get("source_pdb")
        box = int(kwargs["boxsize"])
        if mode != "fixed" or not source or kwargs.get("do_center", False) or box ** 3 * 4 > max_mb * 1024 ** 2:
            return None
        st = os.stat(source)
        return (os.path.realpath(source), st.st_size, st.st_mtime_ns, str(kwargs["chain_id"]), box,
                float(kwargs["apix"]), float(kwargs["res"]), str(kwargs.get("device", "cuda")),
                bool(kwargs.get("include_hetatm", False)), backend)

    def _coordinates(self, pose, key):
        t0 = time.perf_counter()
        moved = transform_coordinates(self.moving_coords, pose)
        t1 = time.perf_counter()
        q = quantize_8_3(moved)[self.transform_rows]
        return dict(key=key, pose=tuple(pose), coordinates=q, transform_s=t1 - t0,
                    quantize_s=time.perf_counter() - t1)

    def coordinate_job(self, kwargs, max_mb=128.0, mode="fixed", backend="auto"):
        key = self.source_key(kwargs, max_mb, mode, backend)
        if key is None or key != self.key or self.fixed is None:
            return None
        pose = tuple(kwargs["pose"])
        return lambda: self._coordinates(pose, key)

    def generate(self, kwargs, max_mb=128.0, mode="fixed", backend="auto", prepared=None):
        from .density import generate_volume
        started = time.perf_counter()
        self.last_metrics = dict(fixed_density_hits=0, fixed_density_cached_atoms=0)
        key = self.source_key(kwargs, max_mb, mode, backend)
        if key is None:
            self.clear()
            args = {k: v for k, v in kwargs.items() if k in ("pdb_path", "boxsize", "apix", "res", "device",
                                                                "do_center", "do_yflip", "include_hetatm",
                                                                "source_pdb", "chain_id", "pose")}
            return generate_volume(**args)
        source, chain = kwargs["source_pdb"], str(kwargs["chain_id"])
        box = int(kwargs["boxsize"])
        include = bool(kwargs.get("include_hetatm", False))
        device = kwargs.get("device", "cuda")
        hit = key == self.key and self.fixed is not None
        if not hit:
            self.clear()
            first = read_pdb_cached(source)
            dens = read_pdb_cached(source, include_hetatm=include, first_model_only=False)
            sel = first.chains == chain
            if not sel.any():
                raise ValueError(f"Chain {chain!r} has no atoms in the first PDB model")
            moving_lines = first.line_numbers[sel]
            self.moving_coords = first.coords[sel]
            positions = {int(line): i for i, line in enumerate(moving_lines)}
            moving = np.isin(dens.line_numbers, moving_lines)
            self.transform_rows = np.asarray([positions[int(line)] for line in dens.line_numbers[moving]],
                                             dtype=np.int64)
            numbers, _ = atoms_to_atomic_numbers_and_masses(dens.elements)
            self.moving_numbers = numbers[moving]
            self.fixed_atom_count = int((~moving).sum())
            self.fixed = density_from_coords(dens.coords[~moving], numbers[~moving], box, kwargs["apix"],
                                             kwargs["res"], device=device)
            self.key = key
        prepared_at = time.perf_counter()
        pose = tuple(kwargs["pose"])
        prefetched = prepared is not None and prepared.get("key") == key and prepared.get("pose") == pose
        if not prefetched:
            prepared = self._coordinates(pose, key)
        t0 = time.perf_counter()
        volume = density_from_coords(prepared["coordinates"], self.moving_numbers, box, kwargs["apix"],
                                     kwargs["res"], device=device, _initial_volume=self.fixed)
        if kwargs.get("do_yflip", False):
            volume = yflip_vectorized(volume)
        self.last_metrics.update(fixed_density_hits=int(hit), fixed_density_cached_atoms=self.fixed_atom_count,
                                 density_moving_atoms=len(self.moving_numbers),
                                 density_prepare_s=prepared_at - started, density_transform_s=prepared["transform_s"],
                                 density_quantize_s=prepared["quantize_s"], density_prefetch_hits=int(prefetched),
                                 density_deposit_s=time.perf_counter() - t0)
        return volume


def search_volumes(engine: SearchEngine, ps: ParticleSet, volumes: Sequence[torch.Tensor], path: str = "fp32",
                   tile: Optional[dict] = None):
    """Banks of several model volumes, then ONE search launch over all of them
    (ZL:gisapr_batch_search.py:251-364, the candidate x direction dimension
    batched).  Returns (per-model float64 scores, CorrelationResult, SlotPlan)."""
    tile = tile or {}
    C, M = len(volumes), len(engine.dir_rot)
    plan = build_slot_plan(ps, M, C)
    if path not in PATHS:
        raise ValueError(f"path must be one of {PATHS}")
    if path == "mixed":
        if ps.coeffs is None:
            engine.prepare_small_shift(ps)
        tk = torch.empty((C * M, engine.small.Kpad), dtype=torch.complex64, device=engine.device)
        for c, vol in enumerate(volumes):
            engine.small_templates(vol, out=tk[c * M:(c + 1) * M])
        corr = kernels.correlate_mixed(tk, ps.coeffs, ps.P, plan.dir_off, plan.dir_idx, engine.small_bhi,
                                       engine.small_blo, engine.small.half, validate=False, pmap=plan.pmap,
                                       pt=tile.get("pt", 0), wpb=tile.get("wpb", 0), plast=engine.small_plast,
                                       operand=engine.small.operand)
    else:
        if ps.spectra_frag is None:
            engine.prepare(ps)
        n = engine.n
        bank = kernels.FragmentPlanes(torch.empty((C * M, kernels.fragment_plane(n)), dtype=torch.complex64,
                                                  device=engine.device), n)
        tinv = torch.empty((C * M,), dtype=torch.float32, device=engine.device)
        for c, vol in enumerate(volumes):
            dft = kernels.volume_dft(vol)
            _, norms = kernels.build_templates(dft, n, engine.rts, engine.mask, engine.fsc_packed,
                                               out=bank[c * M:(c + 1) * M], normalize=False)
            torch.reciprocal(norms, out=tinv[c * M:(c + 1) * M])
        if path == "split":
            corr = kernels.correlate_split(bank, engine.search_spectra(ps), ps.inv_norm, plan.dir_off, plan.dir_idx,
                                           engine.crop, validate=False, pmap=plan.pmap, tinv=tinv,
                                           order=plan.virtual_major(C),
                                           dt=tile.get("dt", 0), pt=tile.get("pt", 0))
        elif path == "split16":
            corr = kernels.correlate_split_f16(bank, engine.search_spectra16(ps), plan.dir_off, plan.dir_idx,
                                               engine.crop, validate=False, pmap=plan.pmap, tinv=tinv,
                                               order=plan.virtual_major(C), dt=tile.get("dt", 0),
                                               pt=tile.get("pt", 0), refine=engine.search_spectra(ps),
                                               inv_norm=ps.inv_norm, bank16=kernels.pack_fragments_f16(bank, tinv))
        else:
            corr = kernels.correlate(bank, engine.search_spectra(ps), ps.inv_norm, plan.dir_off, plan.dir_idx,
                                     engine.crop,
                                     validate=False, pmap=plan.pmap, tinv=tinv, order=plan.virtual_major(C),
                                     dt=tile.get("dt", 0), pt=tile.get("pt", 0))
    scores = kernels.score_sum_segments(corr.peak, corr.best, plan.slot_off)
    return scores, corr, plan


class _ByteCount:
    """Read-only ``bytes`` of what an engine keeps resident."""

    def __init__(self, count: Callable[[], int]):
        self._count = count

    @property
    def bytes(self) -> int:
        return int(self._count())


class BatchSearchEngine:
    """Group evaluation of several models against one particle selection
    (call surface of ZL:gisapr_batch_search.py:142-364).

    ``run(args, models)`` takes the v606 argument namespace and a list of
    ``generate_volume`` keyword dicts (``source_pdb``, ``chain_id``,
    ``pose``, ``boxsize``, ``apix``, ``res``, ...); each model's density
    comes from ``RigidDensityCache``, the particles are read and prepared
    once per context, and every group of models is searched in one launch.
    Returns one dict per model: ``score`` (float64 sum exp(rawCC/1024)),
    ``search_text`` (the three v606 records per particle), ``metrics``.
    ``options``: ``batch_candidates`` (group size), ``path`` ("fp32", the
    reference's search; "split", the same operator on split-bf16 MFMA; or
    "mixed", the small-shift window), ``fixed_density_cache_mb``,
    ``batch_density_mode`` ("fixed" reuses the fixed body).
    ``orientation_rows`` (a calibration bank's rows) restricts every
    particle to its bank directions; ``calibrate`` returns each model's
    per-particle Top-K rows as ``orientation_rows`` (ZL gisapr_batch_search.py
    :207-364 with orientation_bank_v1.py).
    """

    def __init__(self, search_globals=None, device=None, options=None):
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        # the caller's dict as given (the worker compares it to decide whether to
        # keep this engine); defaults fill what it leaves out
        self.options = dict(options or {})
        self._defaults = dict(batch_candidates=8, path="fp32", fixed_density_cache_mb=512.0,
                              batch_density_mode="fixed")
        # resident bytes the worker's ping reports (ZL gisapr_worker_v2.py:181-183)
        self.particles = _ByteCount(lambda: self.context["ps"].nbytes_resident() if self.context else 0)
        self.volumes = _ByteCount(lambda: 0)  # model volumes are not retained between requests
        self.density = RigidDensityCache()
        self.context = None

    def _opt(self, key):
        return self.options.get(key, self._defaults[key])

    def _context(self, args):
        from . import host
        from .search import ParticleMRCReader, read_model_angles, read_particle_table, select_particle_records
        runtime = {"o", "mrc", "i", "gpuid", "verbose"}
        key = tuple(sorted((k, str(v)) for k, v in vars(args).items() if k not in runtime))
        if self.context is not None and self.context["key"] == key:
            return self.context
        N, n = int(args.oriboxsize), int(args.newboxsize)
        angles = read_model_angles(args.ang)
        table = read_particle_table(args.p)
        records = select_particle_records(table, args.start, args.end)
        if not records:
            raise ValueError("No particles selected for batch evaluation")
        engine = SearchEngine(box=N, new_box=n, apix=float(args.apix), crop=int(args.ccgCropSize),
                              psi_step=float(args.psiStep), local_range=float(args.localRange),
                              do_local_search=bool(args.doLocalSearch), kk=float(args.kk),
                              voltage=float(args.voltage), cs=float(args.cs), mask_radius=int(args.maskRadius),
                              mask_edge=int(args.maskEdge), discard_mask=bool(args.discardMask),
                              fsc_values=host.read_fsc_values(args.FSC, N, bool(args.ignoreFSC)),
                              dir_rot=angles.rot, dir_tilt=angles.tilt, device=self.device)
        reader = ParticleMRCReader(N, self.device)
        try:
            images = reader.read_many([r.image_name for r in records])
        finally:
            reader.close()
        dfo = None
        if table.has_ctf:
            dfo = np.asarray([[r.defocus_u, r.defocus_v, r.defocus_angle] for r in records], np.float32)
        ps = engine.add_particles(images, [r.rot for r in records], [r.tilt for r in records],
                                  [r.psi for r in records], defocus=dfo, line_index=[r.line_index for r in records])
        if self._opt("path") == "mixed":
            engine.enable_small_shift()
            engine.prepare_small_shift(ps)
        self.context = dict(key=key, args=args, engine=engine, ps=ps, records=records, angles=angles)
        return self.context

    def _restricted(self, ctx, orientation_rows, keep: int) -> ParticleSet:
        """The context's particle set with each particle's directions cut to
        its bank row (spectra shared)."""
        from dataclasses import replace
        from . import orientation_bank
        ps, angles = ctx["ps"], ctx["angles"]
        local = [orientation_bank.restricted_indices(orientation_rows.get(str(rec.line_index)), rec, ps.local[b],
                                                     angles, keep)
                 for b, rec in enumerate(ctx["records"])]
        off = np.zeros(ps.B + 1, np.int32)
        off[1:] = np.cumsum([len(l) for l in local])
        idx = np.asarray([k for l in local for k in l], np.int32)
        return replace(ps, local=local, dir_off=torch.from_numpy(off).to(ps.dir_off.device),
                       dir_idx=torch.from_numpy(idx).to(ps.dir_off.device), row_order=None)

    def run(self, args, models, orientation_rows=None, calibrate=False):
        from . import orientation_bank
        from .search import _records
        if calibrate and orientation_rows is not None:
            raise ValueError("Cannot calibrate from restricted directions")
        keep = int(getattr(args, "keep_orientations", 3))
        t0 = time.perf_counter()
        ctx = self._context(args)
        setup_s = time.perf_counter() - t0
        engine, ps, records, angles = ctx["engine"], ctx["ps"], ctx["records"], ctx["angles"]
        if orientation_rows is not None:
            ps = self._restricted(ctx, orientation_rows, keep)
        path = self._opt("path")
        center = engine.small.half if path == "mixed" else engine.crop // 2
        P, M = ps.P, len(angles.rot)
        results = []
        cap = max(1, int(self._opt("batch_candidates")))
        for g0 in range(0, len(models), cap):
            group = models[g0:g0 + cap]
            t1 = time.perf_counter()
            volumes = [self.density.generate(dict(m, device=str(self.device)),
                                             self._opt("fixed_density_cache_mb"),
                                             self._opt("batch_density_mode")) for m in group]
            torch.cuda.synchronize(self.device)
            t2 = time.perf_counter()
            scores, corr, plan = search_volumes(engine, ps, volumes, path)
            sc = scores.cpu().numpy()
            torch.cuda.synchronize(self.device)
            t3 = time.perf_counter()
            best = corr.best.cpu().numpy()
            peak = corr.peak.cpu().numpy()
            mean = corr.mean.cpu().numpy()
            std = corr.std.cpu().numpy()
            py = corr.peak_y.cpu().numpy().reshape(-1)
            px = corr.peak_x.cpu().numpy().reshape(-1)
            d_off = plan.dir_off.cpu().numpy()
            d_idx = plan.dir_idx.cpu().numpy()
            all_scores = corr.scores.reshape(-1, P).cpu().numpy() if calibrate else None
            for c in range(len(group)):
                lines = []
                ranked = {} if calibrate else None
                for j in range(int(plan.slot_off_host[c]), int(plan.slot_off_host[c + 1])):
                    b = j - int(plan.slot_off_host[c])
                    if calibrate:
                        ranked[str(records[b].line_index)] = orientation_bank.ranked_row(
                            records[b], ps.local[b], angles, all_scores[d_off[j]:d_off[j + 1]], keep)
                    if best[j] < 0:
                        continue
                    li, pi = divmod(int(best[j]), P)
                    g = (d_off[j] + li) * P + pi
                    gi = int(d_idx[d_off[j] + li]) - c * M
                    for item in _records(records[b], ps.fake_psi[b], angles.rot, angles.tilt, gi, pi,
                                         float(peak[j]), int(py[g]), int(px[g]), float(mean[j]), float(std[j]),
                                         2 * center, engine.trans_scale):
                        lines.append(str(item) + "\n")
                n_g = len(group)
                metrics = dict(density_s=(t2 - t1) / n_g, search_s=(t3 - t2) / n_g, setup_s=setup_s / len(models),
                               **self.density.last_metrics)
                metrics["execution_s"] = metrics["density_s"] + metrics["search_s"] + metrics["setup_s"]
                results.append(dict(score=float(sc[c]), orientation_rows=ranked, search_text="".join(lines),
                                    metrics=metrics))
        return results
