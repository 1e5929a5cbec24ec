"""Device-resident candidate scoring: density -> template bank -> search -> score.

``SearchEngine`` holds everything that is fixed per dataset on the HIP
device (mask, packed FSC, direction matrices, CTF tables, correlation
tables) and ``ParticleSet`` holds one particle stack's spectra, psi grids and
local direction lists resident in HBM, so a candidate evaluation only runs
the splat, the 3-D transform, the template bank, the correlation and the
per-particle reduction.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import host, kernels
from .density import splat_parameters


@dataclass
class ParticleSet:
    images: torch.Tensor               # [B, N, N] float32 (device)
    rot: np.ndarray
    tilt: np.ndarray
    psi: np.ndarray
    line_index: np.ndarray
    local: List[List[int]]             # per particle: global direction indices
    fake_psi: np.ndarray               # [B, P] searched psi (record value)
    eff_psi: np.ndarray                # [B, P] effective psi (0.5 deg quirk)
    rotp: torch.Tensor                 # [B, P, 6]
    defocus: Optional[torch.Tensor]    # [B, 3] or None
    dir_off: torch.Tensor              # [B+1] int32
    dir_idx: torch.Tensor              # [sum M_i] int32
    spectra: Optional[torch.Tensor] = None   # [B, P, n, n/2+1] complex64 row-major copy (on demand)
    inv_norm: Optional[torch.Tensor] = None  # [B, P] float32
    coeffs: Optional[torch.Tensor] = None    # [B*P, Kpad] complex64, unit norm (small-shift search)
    row_order: Optional[torch.Tensor] = None # [sum M_i] int32 processing order of the small-shift search
    particle_order: Optional[torch.Tensor] = None  # [B] int32 processing order of the split search
    spectra_frag: Optional[kernels.FragmentPlanes] = None  # [B, P] psi spectra in fragment order (searched)
    spectra16: Optional[kernels.HalfFragmentPlanes] = None  # [B, P] unit-norm x 2^7 fp16 spectra (fp16 search)

    @property
    def B(self) -> int:
        return int(self.images.shape[0])

    @property
    def P(self) -> int:
        return int(self.rotp.shape[1])

    @property
    def n_scores(self) -> int:
        return int(self.dir_idx.shape[0]) * self.P

    def nbytes_resident(self) -> int:
        t = [self.images, self.rotp, self.dir_off, self.dir_idx, self.spectra, self.inv_norm, self.defocus,
             self.coeffs, None if self.spectra_frag is None else self.spectra_frag.data,
             None if self.spectra16 is None else self.spectra16.data]
        return sum(x.numel() * x.element_size() for x in t if x is not None)


def search_bank(bank) -> kernels.FragmentPlanes:
    """A template bank in the searches' operand layout: FragmentPlanes as
    they are, a row-major [M, n, ld] bank packed (one permuting copy)."""
    return bank if isinstance(bank, kernels.FragmentPlanes) else kernels.pack_fragments(bank.contiguous())


def particle_order(dir_off: np.ndarray, dir_idx: np.ndarray, device) -> torch.Tensor:
    """Particles sorted by their first local direction (int32 [B] on ``device``):
    neighbours in this order share most of their template rows."""
    off = np.asarray(dir_off, np.int64)
    idx = np.asarray(dir_idx, np.int64)
    first = np.where(off[1:] > off[:-1], idx[np.minimum(off[:-1], max(len(idx) - 1, 0))], -1)
    return torch.from_numpy(np.argsort(first, kind="stable").astype(np.int32)).to(device)


@dataclass
class SearchResult:
    corr: kernels.CorrelationResult
    score: torch.Tensor  # [1] float64: sum exp(rawCC / 1024)
    center: int = 0      # peak index of the zero lag (crop // 2, or half for the small-shift search)


class SearchEngine:
    """One search configuration (box, crop, psi sampling, CTF, mask, FSC, directions)."""

    def __init__(self, *, box, new_box=None, apix, crop, psi_step, local_range=20.0,
                 do_local_search=True, kk=0.0, voltage=300.0, cs=2.7, mask_radius=110, mask_edge=6,
                 discard_mask=False, fsc_values=None, dir_rot, dir_tilt, device=None):
        self.N = int(box)
        self.n = int(self.N if new_box is None else new_box)
        if self.N % 2 or self.n % 2:
            raise ValueError("oriboxsize and newboxsize must be even")
        if self.n > self.N:
            raise ValueError("newboxsize cannot exceed oriboxsize")
        if psi_step <= 0:
            raise ValueError("psiStep must be positive")
        if crop <= 0 or crop % 2 or crop > self.n:
            raise ValueError("CCG crop size must be positive, even and <= newboxsize")
        if crop > 32:
            raise ValueError(f"--ccgCropSize {crop}: the searches cover crops of up to 32 x 32 lags")
        if self.N not in kernels.SUPPORTED_BOXES:
            raise ValueError(f"box {self.N} is not one of {kernels.SUPPORTED_BOXES}")
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        if self.device.type != "cuda":
            raise ValueError("gisapr_amd runs on HIP devices only")
        self.apix = float(apix)
        self.new_apix = self.N / self.n * self.apix
        self.trans_scale = self.apix / self.new_apix
        self.crop = int(crop)
        self.psi_step = float(psi_step)
        self.local_range = float(local_range)
        self.do_local_search = bool(do_local_search)
        self.dir_rot = np.asarray(dir_rot, np.float64)
        self.dir_tilt = np.asarray(dir_tilt, np.float64)
        dev = self.device
        self.mask = None if discard_mask else torch.from_numpy(
            host.generate_soft_mask((self.N, self.N), int(mask_radius), int(mask_edge))).to(dev)
        fsc = np.ones(self.N, np.float32) if fsc_values is None else np.asarray(fsc_values, np.float32)
        self.fsc_packed = torch.from_numpy(host.packed_fsc(fsc, self.n)).to(dev)
        self.rts = torch.from_numpy(np.ascontiguousarray(host.rotation_matrices_t(self.dir_rot, self.dir_tilt))).to(dev)
        self.ctf = host.CTFTables(self.n, self.new_apix, cs, voltage, kk)
        self.ctf_tab = torch.from_numpy(self.ctf.table).to(dev)

    # ---------------------------------------------------------------- particles
    def local_directions(self, rot, tilt) -> List[List[int]]:
        """Each particle's searched directions: the local map (v606:351-378)
        with --doLocalSearch, else every direction."""
        rot, tilt = np.asarray(rot, np.float64).reshape(-1), np.asarray(tilt, np.float64).reshape(-1)
        if self.do_local_search:
            return host.local_orientation_map(rot, tilt, self.dir_rot, self.dir_tilt, self.local_range)
        return [list(range(len(self.dir_rot))) for _ in range(len(rot))]

    def add_particles(self, images, rot, tilt, psi, defocus=None, line_index=None,
                      prepare=True, local: Optional[Sequence[Sequence[int]]] = None) -> ParticleSet:
        """``local`` overrides the searched directions per particle (ascending
        global indices), e.g. a calibration bank's Top-K rows."""
        dev = self.device
        images = torch.as_tensor(images, dtype=torch.float32).to(dev).contiguous()
        B = int(images.shape[0])
        if tuple(images.shape[1:]) != (self.N, self.N):
            raise ValueError(f"particle images must be [B, {self.N}, {self.N}]")
        rot, tilt, psi = (np.asarray(v, np.float64).reshape(B) for v in (rot, tilt, psi))
        if local is None:
            local = self.local_directions(rot, tilt)
        else:
            local = [sorted(int(i) for i in l) for l in local]
            if len(local) != B or any(i < 0 or i >= len(self.dir_rot) for l in local for i in l):
                raise ValueError("local must list valid direction indices for every particle")
        fakes, effs = [], []
        for p in psi:
            f, e = host.psi_sampling_for_record(float(p), self.psi_step, self.do_local_search, self.local_range)
            fakes.append(f)
            effs.append(e)
        fake = np.asarray(fakes, np.float64)
        eff = np.asarray(effs, np.float64)
        rotp = np.stack([host.psi_grid_params(self.N, e) for e in effs]).astype(np.float32)
        off = np.zeros(B + 1, np.int32)
        off[1:] = np.cumsum([len(l) for l in local])
        idx = np.asarray([k for l in local for k in l], np.int32)
        dfo = None
        if defocus is not None:
            dfo = torch.from_numpy(np.ascontiguousarray(np.asarray(defocus, np.float32).reshape(B, 3))).to(dev)
        ps = ParticleSet(
            images=images, rot=rot, tilt=tilt, psi=psi,
            line_index=np.arange(B) if line_index is None else np.asarray(line_index),
            local=local, fake_psi=fake, eff_psi=eff,
            rotp=torch.from_numpy(rotp).to(dev), defocus=dfo,
            dir_off=torch.from_numpy(off).to(dev),
            dir_idx=torch.from_numpy(idx if len(idx) else np.zeros(0, np.int32)).to(dev))
        if prepare:
            self.prepare(ps)
        return ps

    def prepare(self, ps: ParticleSet) -> ParticleSet:
        """Psi spectra in fragment order (the searches' operand layout) and
        their 1/norms; a row-major copy made earlier (``row_spectra``) is
        refreshed."""
        ctf_tab = self.ctf_tab if ps.defocus is not None else None
        if ps.spectra_frag is None:
            ps.spectra_frag = kernels.FragmentPlanes(
                torch.empty((ps.B, ps.P, kernels.fragment_plane(self.n)), dtype=torch.complex64,
                            device=self.device), self.n)
        _, ps.inv_norm = kernels.prepare_particles(
            ps.images, self.n, ps.rotp, self.mask, ctf_tab, ps.defocus, self.ctf.kconst,
            out=ps.spectra_frag, inv_norm=ps.inv_norm)
        if ps.spectra is not None:
            ps.spectra.copy_(ps.spectra_frag.unpack())
        if ps.spectra16 is not None:
            kernels.pack_fragments_f16(ps.spectra_frag, ps.inv_norm, out=ps.spectra16.data)
        return ps

    def row_spectra(self, ps: ParticleSet) -> torch.Tensor:
        """The psi spectra row major [B, P, n, n/2+1] (unpacked from the
        fragment-order planes once, then kept in step by ``prepare``)."""
        if ps.spectra_frag is None:
            self.prepare(ps)
        if ps.spectra is None:
            ps.spectra = ps.spectra_frag.unpack()
        return ps.spectra

    # ---------------------------------------------------------------- model
    def density(self, coords, numbers, res, base: Optional[torch.Tensor] = None) -> torch.Tensor:
        width, rp, kn = splat_parameters(self.apix, res)
        xyz = torch.as_tensor(np.asarray(coords, np.float32)).to(self.device).contiguous()
        zn = torch.as_tensor(np.asarray(numbers, np.float32)).to(self.device).contiguous()
        vol = base.clone() if base is not None else None
        return kernels.splat(xyz, zn, self.N, width, self.apix, rp, kn, vol)

    def templates(self, volume: torch.Tensor, with_norms=False, normalize=True, fragments=False):
        """Unit-norm bank (and norms); with ``normalize=False`` returns
        (bank, 1/norms) with unnormalised rows, for ``search(..., tinv=)``.
        ``fragments``: the bank as FragmentPlanes, the searches' operand
        layout, written by the template kernel directly."""
        dft = kernels.volume_dft(volume)
        out = None
        if fragments:
            out = kernels.FragmentPlanes(torch.empty((len(self.dir_rot), kernels.fragment_plane(self.n)),
                                                     dtype=torch.complex64, device=self.device), self.n)
        out = kernels.build_templates(dft, self.n, self.rts, self.mask, self.fsc_packed, with_norms, out=out,
                                      normalize=normalize)
        if normalize:
            return out
        bank, norms = out
        return bank, 1.0 / norms

    # ---------------------------------------------------------------- search
    def search(self, bank: torch.Tensor, ps: ParticleSet, tinv: Optional[torch.Tensor] = None) -> SearchResult:
        if ps.spectra_frag is None:
            self.prepare(ps)
        if int(bank.shape[0]) != len(self.dir_rot):
            raise ValueError("template bank does not cover the engine's direction table")
        # dir_off / dir_idx were built on the host against this direction table
        if ps.particle_order is None:
            ps.particle_order = particle_order(ps.dir_off.cpu().numpy(), ps.dir_idx.cpu().numpy(), self.device)
        corr = kernels.correlate(search_bank(bank), self.search_spectra(ps), ps.inv_norm, ps.dir_off, ps.dir_idx,
                                 self.crop, validate=False, tinv=tinv, order=ps.particle_order)
        return SearchResult(corr=corr, score=kernels.score_sum(corr.peak, corr.best), center=self.crop // 2)

    def search_spectra(self, ps: ParticleSet) -> kernels.FragmentPlanes:
        """The particle spectra in fragment order, the searches' operand
        layout (written by ``prepare``)."""
        if ps.spectra_frag is None:
            self.prepare(ps)
        return ps.spectra_frag

    def search_spectra16(self, ps: ParticleSet) -> kernels.HalfFragmentPlanes:
        """The particle spectra as the fp16-operand search reads them (unit
        norm x 2^7, fp16), made from the fragment-order spectra once and kept
        in step by ``prepare``."""
        if ps.spectra_frag is None:
            self.prepare(ps)
        if ps.spectra16 is None:
            ps.spectra16 = kernels.pack_fragments_f16(ps.spectra_frag, ps.inv_norm)
        return ps.spectra16

    def search_split(self, bank: torch.Tensor, ps: ParticleSet, tinv: Optional[torch.Tensor] = None,
                     dt: int = 0, pt: int = 0, operand: str = "bf16", wsplit: bool = False,
                     refine: bool = True) -> SearchResult:
        """The search of ``search`` (every coefficient, every lag of the crop)
        on split-bf16 MFMA; particles are processed in orientation order.
        ``operand="fp16"``: fp16 operands and packed fp16 products (the row
        table split into fp16 hi + lo with ``wsplit``), then, with ``refine``,
        each particle's best pair and its near ties rescored on split-bf16."""
        if operand not in ("bf16", "fp16"):
            raise ValueError("operand must be 'bf16' or 'fp16'")
        if ps.spectra_frag is None:
            self.prepare(ps)
        if int(bank.shape[0]) != len(self.dir_rot):
            raise ValueError("template bank does not cover the engine's direction table")
        if ps.particle_order is None:
            ps.particle_order = particle_order(ps.dir_off.cpu().numpy(), ps.dir_idx.cpu().numpy(), self.device)
        if operand == "fp16":
            fb = search_bank(bank)
            corr = kernels.correlate_split_f16(fb, self.search_spectra16(ps), ps.dir_off, ps.dir_idx, self.crop,
                                               validate=False, tinv=tinv, order=ps.particle_order, dt=dt, pt=pt,
                                               wsplit=wsplit, refine=ps.spectra_frag if refine else None,
                                               inv_norm=ps.inv_norm, bank16=kernels.pack_fragments_f16(fb, tinv))
        else:
            corr = kernels.correlate_split(search_bank(bank), self.search_spectra(ps), ps.inv_norm, ps.dir_off,
                                           ps.dir_idx, self.crop, validate=False, tinv=tinv,
                                           order=ps.particle_order, dt=dt, pt=pt)
        return SearchResult(corr=corr, score=kernels.score_sum(corr.peak, corr.best), center=self.crop // 2)

    # ---------------------------------------------------------------- small-shift precision search
    def enable_small_shift(self, resolution: float = 4.0, half: int = 3,
                           operand: str = "bf16") -> host.SmallShiftTables:
        """Set up the split-bf16 (``operand="fp16"``: split-fp16) search over
        lags [-half, half]^2 and the coefficients with |k| <= N apix /
        resolution (capped at n/2)."""
        radius = min(self.N * self.apix / float(resolution), self.n / 2.0)
        return self._install_small(host.SmallShiftTables(self.n, radius, half, operand))

    def set_small_shift_operand(self, operand: str) -> host.SmallShiftTables:
        """Switch the small-shift search between split-bf16 and split-fp16
        operands, keeping its coefficient set and lags (prepared particle
        coefficients stay valid)."""
        self._require_small()
        return self._install_small(host.SmallShiftTables(self.n, self.small.radius, self.small.half, operand))

    def _install_small(self, tab: host.SmallShiftTables) -> host.SmallShiftTables:
        dev = self.device
        self.small = tab
        self.small_kidx = torch.from_numpy(tab.kidx).to(dev)
        self.small_kslot = torch.from_numpy(tab.kslot).to(dev)
        self.small_bhi = torch.from_numpy(tab.b_hi.view(np.int16)).to(dev)
        self.small_blo = torch.from_numpy(tab.b_lo.view(np.int16)).to(dev)
        self.small_plast = torch.from_numpy(tab.phi_last).to(dev)
        return tab

    def _require_small(self):
        if getattr(self, "small", None) is None:
            raise RuntimeError("call enable_small_shift() first")

    def prepare_small_shift(self, ps: ParticleSet, drop_spectra: bool = False) -> ParticleSet:
        """Gather the unit-norm in-disk coefficients of every psi spectrum (once per dataset)."""
        self._require_small()
        ps.coeffs = kernels.gather_coefficients(self.row_spectra(ps), self.small_kidx, self.small.Kpad, ps.inv_norm)
        if drop_spectra:
            ps.spectra = ps.spectra_frag = None
        return ps

    def template_coefficients(self, bank: torch.Tensor, tinv: Optional[torch.Tensor] = None) -> torch.Tensor:
        """The in-disk coefficients of a bank (times 1/norm per row when given)."""
        self._require_small()
        return kernels.gather_coefficients(bank, self.small_kidx, self.small.Kpad, tinv)

    def small_templates(self, volume: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        """[M, Kpad] unit-norm template coefficients of a model volume, built
        without storing the bank (the small-shift search's template operand)."""
        self._require_small()
        dft = kernels.volume_dft(volume)
        tk, _ = kernels.build_template_coefficients(dft, self.n, self.rts, self.mask, self.fsc_packed,
                                                    self.small_kslot, self.small.K, self.small.Kpad, out=out)
        return tk

    def search_small_shift(self, bank: Optional[torch.Tensor], ps: ParticleSet,
                           tinv: Optional[torch.Tensor] = None, tk: Optional[torch.Tensor] = None) -> SearchResult:
        """Best lag in [-half, half]^2 per (direction, psi) on split-bf16 MFMA,
        from a bank (gathered here) or from template coefficients ``tk``."""
        self._require_small()
        if ps.coeffs is None:
            self.prepare_small_shift(ps)
        if tk is None:
            if int(bank.shape[0]) != len(self.dir_rot):
                raise ValueError("template bank does not cover the engine's direction table")
            tk = self.template_coefficients(bank, tinv)
        elif int(tk.shape[0]) != len(self.dir_rot):
            raise ValueError("template coefficients do not cover the engine's direction table")
        if getattr(ps, "row_order", None) is None:
            from .batch import orientation_rank
            off = ps.dir_off.cpu().numpy()
            rank = orientation_rank(off, ps.dir_idx.cpu().numpy())
            row_particle = np.repeat(np.arange(ps.B), np.diff(off))
            ps.row_order = torch.from_numpy(np.argsort(rank[row_particle], kind="stable").astype(np.int32)).to(
                self.device)
        corr = kernels.correlate_mixed(tk, ps.coeffs, ps.P, ps.dir_off, ps.dir_idx, self.small_bhi,
                                       self.small_blo, self.small.half, validate=False, plast=self.small_plast,
                                       rows=ps.row_order, operand=self.small.operand)
        return SearchResult(corr=corr, score=kernels.score_sum(corr.peak, corr.best), center=self.small.half)

    def score_volume(self, volume: torch.Tensor, ps: ParticleSet) -> SearchResult:
        return self.search(self.templates(volume), ps)

    # ---------------------------------------------------------------- records
    def records(self, res: SearchResult, ps: ParticleSet) -> List[dict]:
        """Per-particle results (make_result_records fields, v606:1452-1516)."""
        c = res.corr
        best = c.best.cpu().numpy()
        peak = c.peak.cpu().numpy()
        mean = c.mean.cpu().numpy()
        std = c.std.cpu().numpy()
        py = c.peak_y.cpu().numpy()
        px = c.peak_x.cpu().numpy()
        off = ps.dir_off.cpu().numpy()
        idx = ps.dir_idx.cpu().numpy()
        P = ps.P
        out = []
        for b in range(ps.B):
            if best[b] < 0:
                continue
            li, pi = divmod(int(best[b]), P)
            g = (off[b] + li) * P + pi
            gi = int(idx[off[b] + li])
            ad = host.calculate_angular_distance(self.dir_rot[gi], self.dir_tilt[gi], 0.0,
                                                 float(ps.rot[b]), float(ps.tilt[b]), 0.0)
            z = (float(peak[b]) - mean[b]) / std[b] if std[b] > 0 else 0.0
            yy, xx = int(py.reshape(-1)[g]), int(px.reshape(-1)[g])
            out.append({
                "line": int(ps.line_index[b]), "z": float(z), "l3": 1 if z > 3.0 else 0,
                "accu": 1 if ad < 11.0 else 0, "rot": float(self.dir_rot[gi]), "tilt": float(self.dir_tilt[gi]),
                "angdist": ad,
                "xshift": round(float(res.center - xx) / self.trans_scale),
                "yshift": round(float(res.center - yy) / self.trans_scale),
                "psi_fake": float(ps.fake_psi[b, pi]), "psi_prior": float((ps.psi[b] + 360.0) % 360.0),
                "rawcc": float(peak[b]), "best_local": li, "best_psi": pi, "peak_y": yy, "peak_x": xx,
                "direction": gi, "mean": float(mean[b]), "std": float(std[b]),
            })
        return out
