"""Typed wrappers of the C ABI for torch tensors on a HIP device.

Each function validates shapes on the host, allocates outputs and workspace
with the torch caching allocator, and launches on the current torch stream.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from ._lib import check, lib, ptr, require_hip, stream_of

SUPPORTED_BOXES = (32, 64, 96, 128, 192, 256)


def _dev(t):
    return t.device


def splat(xyz: torch.Tensor, numbers: torch.Tensor, box: int, width: int, apix: float, rp: float,
          kn: float, volume: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Add the radial-exponential splat of the atoms into ``volume`` ([box]^3, zeroed if None)."""
    require_hip(xyz, "xyz")
    require_hip(numbers, "numbers")
    if xyz.dtype != torch.float32 or numbers.dtype != torch.float32:
        raise ValueError("xyz and numbers must be float32")
    if xyz.ndim != 2 or xyz.shape[1] != 3 or numbers.shape != (xyz.shape[0],):
        raise ValueError("xyz must be [n, 3] and numbers [n]")
    if volume is None:
        volume = torch.zeros((box, box, box), dtype=torch.float32, device=xyz.device)
    require_hip(volume, "volume")
    if tuple(volume.shape) != (box, box, box) or volume.dtype != torch.float32:
        raise ValueError("volume must be float32 [box, box, box]")
    check(lib().gis_splat(ptr(xyz), ptr(numbers), int(xyz.shape[0]), int(box), int(width), float(apix),
                          float(rp), float(kn), ptr(volume), stream_of(xyz.device)), "gis_splat")
    return volume


def move_chain(xyz: torch.Tensor, centre, rt: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Coordinates [k, n, 3] float32 of the chain ``xyz`` ([n, 3] float32) moved
    by each pose and 8.3-rounded: ``rt`` [k, 12] float64 holds R row-major
    ((x - centre) R) and t; ``centre`` is the chain's float32 mean (3 values)."""
    require_hip(xyz, "xyz")
    require_hip(rt, "rt")
    if xyz.dtype != torch.float32 or xyz.ndim != 2 or xyz.shape[1] != 3 or not xyz.is_contiguous():
        raise ValueError("xyz must be contiguous float32 [n, 3]")
    if rt.dtype != torch.float64 or rt.ndim != 2 or rt.shape[1] != 12 or not rt.is_contiguous():
        raise ValueError("rt must be contiguous float64 [k, 12]")
    c = np.ascontiguousarray(np.asarray(centre, np.float32).reshape(3))
    k, n = int(rt.shape[0]), int(xyz.shape[0])
    if out is None:
        out = torch.empty((k, n, 3), dtype=torch.float32, device=xyz.device)
    require_hip(out, "out")
    if tuple(out.shape) != (k, n, 3) or out.dtype != torch.float32 or not out.is_contiguous():
        raise ValueError("out must be contiguous float32 [k, n, 3]")
    check(lib().gis_move_chain(ptr(xyz), n, c.ctypes.data, ptr(rt), k, ptr(out), stream_of(xyz.device)),
          "gis_move_chain")
    return out


def volume_dft(volume: torch.Tensor, out: Optional[torch.Tensor] = None,
               workspace: Optional[torch.Tensor] = None, pad: bool = True) -> torch.Tensor:
    """[N]^3 float32 -> bricked half spectrum of [L, L, L/2+1] complex64
    (pad to L = 2N or not, sinc^2, shift, rfftn); ``volume_dft_natural`` unpacks it."""
    require_hip(volume, "volume")
    N = int(volume.shape[0])
    if volume.ndim != 3 or len(set(volume.shape)) != 1 or volume.dtype != torch.float32:
        raise ValueError("volume must be a cubic float32 tensor")
    pad = int(bool(pad))
    nb = int(lib().gis_volume_dft_bytes_ex(N, pad)) // 8
    if out is None:
        out = torch.empty((nb,), dtype=torch.complex64, device=volume.device)
    elif out.numel() != nb or out.dtype != torch.complex64:
        raise ValueError(f"out must be {nb} complex64 values")
    wsb = int(lib().gis_volume_dft_workspace_ex(N, pad))
    if workspace is not None:
        require_hip(workspace, "workspace")
        if workspace.numel() * workspace.element_size() < wsb:
            raise ValueError(f"workspace must hold {wsb} bytes")
        ws = workspace
    else:
        ws = torch.empty((max(wsb, 1),), dtype=torch.uint8, device=volume.device)
    check(lib().gis_volume_dft_ex(ptr(volume), N, pad, ptr(out), ptr(ws), wsb, stream_of(volume.device)),
          "gis_volume_dft")
    return out


def volume_dft_natural(dft: torch.Tensor, N: int, pad: bool = True) -> torch.Tensor:
    """Unpack the bricked spectrum to natural order [L, L, L/2+1] (for inspection and tests)."""
    L = 2 * N if pad else N
    xb = L // 8 + 1
    b = dft.view(L // 2, L // 2, xb, 2, 2, 4).permute(0, 3, 1, 4, 2, 5).reshape(L, L, 4 * xb)
    return b[..., :L // 2 + 1].contiguous()


def _spectrum_box(dft: torch.Tensor, pad: bool) -> int:
    N = next((b for b in SUPPORTED_BOXES if int(lib().gis_volume_dft_bytes_ex(b, int(pad))) == dft.numel() * 8), 0)
    if not N or dft.dtype != torch.complex64:
        raise ValueError("dft must be a bricked spectrum from volume_dft (with the same pad)")
    return N


def project_slices(dft: torch.Tensor, rts: torch.Tensor, whiten: bool = False, normalize: bool = False,
                   eps: float = 1e-12, pad: bool = True, out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """[M, N, N] float32 real-space central-slice projections of the volume
    whose spectrum is ``dft``; optionally ring-divided (whiten) and plane
    normalised (normalize)."""
    require_hip(dft, "dft")
    require_hip(rts, "rts")
    N = _spectrum_box(dft, pad)
    M = int(rts.shape[0])
    if tuple(rts.shape[1:]) != (3, 3) or rts.dtype != torch.float32 or not rts.is_contiguous():
        raise ValueError("rts must be contiguous float32 [M, 3, 3]")
    if out is None:
        out = torch.empty((M, N, N), dtype=torch.float32, device=dft.device)
    elif tuple(out.shape) != (M, N, N) or out.dtype != torch.float32 or not out.is_contiguous():
        raise ValueError("out must be contiguous float32 [M, N, N]")
    wsb = int(lib().gis_build_templates_workspace_ex(N, M, int(bool(pad))))
    ws = torch.empty((max(wsb, 1),), dtype=torch.uint8, device=dft.device)
    check(lib().gis_project_slices(ptr(dft), N, int(bool(pad)), ptr(rts), M, int(bool(whiten)), int(bool(normalize)),
                                   float(eps), ptr(out), ptr(ws), wsb, stream_of(dft.device)), "gis_project_slices")
    return out


def natural_to_bricked(dft: torch.Tensor) -> torch.Tensor:
    """A natural-order [L, L, L/2+1] complex64 half spectrum (rows and planes
    in natural frequency order) rearranged into the bricked layout."""
    L = int(dft.shape[0])
    if tuple(dft.shape) != (L, L, L // 2 + 1) or L % 8 or dft.dtype != torch.complex64:
        raise ValueError("dft must be complex64 [L, L, L/2+1], L a multiple of 8")
    xb = L // 8 + 1
    padded = torch.zeros((L, L, 4 * xb), dtype=dft.dtype, device=dft.device)
    padded[..., :L // 2 + 1] = dft
    return padded.view(L // 2, 2, L // 2, 2, xb, 4).permute(0, 2, 4, 1, 3, 5).contiguous().view(-1)


def extract_slices(dft_bricked: torch.Tensor, L: int, rts: torch.Tensor) -> torch.Tensor:
    """[M, L, L/2+1] complex64 central slices (rows fftshifted) sampled from a
    bricked half spectrum of transform length L."""
    require_hip(dft_bricked, "dft")
    require_hip(rts, "rts")
    L = int(L)
    if dft_bricked.dtype != torch.complex64 or dft_bricked.numel() != L * L * (L // 8 + 1) * 4:
        raise ValueError("dft must be a bricked complex64 spectrum of length L")
    M = int(rts.shape[0])
    if tuple(rts.shape[1:]) != (3, 3) or rts.dtype != torch.float32 or not rts.is_contiguous():
        raise ValueError("rts must be contiguous float32 [M, 3, 3]")
    wsb = int(lib().gis_extract_slices_workspace(L, max(M, 1)))
    if wsb == 0:
        raise ValueError(f"transform length {L} is not a supported box or twice one")
    out = torch.empty((M, L, L // 2 + 1), dtype=torch.complex64, device=dft_bricked.device)
    ws = torch.empty((wsb,), dtype=torch.uint8, device=dft_bricked.device)
    check(lib().gis_extract_slices(ptr(dft_bricked), L, ptr(rts), M, ptr(out), ptr(ws), wsb,
                                   stream_of(dft_bricked.device)), "gis_extract_slices")
    return out


def sinc2_filter_(volume: torch.Tensor) -> torch.Tensor:
    """volume *= sinc^2(|f|) in place (f = fftshift(fftfreq) per axis) for a
    contiguous float32 [D, H, W] volume; returns the volume."""
    require_hip(volume, "volume")
    if volume.ndim != 3:
        raise ValueError("volume must be 3-D")
    if volume.dtype != torch.float32 or not volume.is_contiguous():
        raise ValueError("volume must be a contiguous float32 tensor")
    if volume.numel():
        D, H, W = (int(v) for v in volume.shape)
        check(lib().gis_sinc2_filter(ptr(volume), D, H, W, stream_of(volume.device)), "gis_sinc2_filter")
    return volume


def slice_sample_grid(base_zyx: torch.Tensor, rts: torch.Tensor, image_shape, dft_shape):
    """grid_sample coordinates [M, *base, 3] (x, y, z order, align_corners=True)
    of M rotations of the base frequency points ``base_zyx`` [..., 3] (z, y, x)
    in an rfft volume of shape ``dft_shape`` taken from a real image of shape
    ``image_shape``, and the conjugation mask [M, *base] (bool)."""
    require_hip(base_zyx, "base_zyx")
    require_hip(rts, "rotation_matrices")
    if base_zyx.dtype != torch.float32 or base_zyx.shape[-1] != 3:
        raise ValueError("base grid must be float32 [..., 3]")
    if rts.dtype != torch.float32 or rts.ndim != 3 or tuple(rts.shape[1:]) != (3, 3):
        raise ValueError("rotation matrices must be float32 [M, 3, 3]")
    img = tuple(int(v) for v in image_shape)
    dsh = tuple(int(v) for v in dft_shape)
    if len(img) != 3 or len(dsh) != 3:
        raise ValueError("image_shape and dft_shape must have three dimensions")
    lead = tuple(base_zyx.shape[:-1])
    M, P = int(rts.shape[0]), int(base_zyx.numel() // 3)
    grid = torch.empty((M, *lead, 3), dtype=torch.float32, device=base_zyx.device)
    conj = torch.empty((M, *lead), dtype=torch.bool, device=base_zyx.device)
    if M and P:
        base = base_zyx.contiguous()
        R = rts.contiguous()
        check(lib().gis_slice_sample_grid(ptr(base), P, ptr(R), M, *img, *dsh, ptr(grid), ptr(conj),
                                          stream_of(base.device)), "gis_slice_sample_grid")
    return grid, conj


def sample_dft_3d(dft: torch.Tensor, coordinates: torch.Tensor) -> torch.Tensor:
    """Trilinear, border-clamped samples [...] complex64 of a complex64 [D, H, W]
    volume at array coordinates [..., 3] (z, y, x)."""
    require_hip(dft, "dft")
    if dft.dtype != torch.complex64 or dft.ndim != 3:
        raise ValueError("dft must be a complex64 [D, H, W] volume")
    if not isinstance(coordinates, torch.Tensor) or coordinates.shape[-1] != 3:
        raise ValueError("coordinates must be [..., 3] (z, y, x)")
    c = coordinates.to(dft.device, torch.float32).contiguous()
    lead = tuple(c.shape[:-1])
    out = torch.empty(lead, dtype=torch.complex64, device=dft.device)
    P = int(out.numel())
    if P:
        D, H, W = (int(v) for v in dft.shape)
        check(lib().gis_sample_dft_3d(ptr(dft), D, H, W, ptr(c), P, ptr(out), stream_of(dft.device)),
              "gis_sample_dft_3d")
    return out


def whiten_images(images: torch.Tensor, whiten: bool = True, normalize: bool = True, eps: float = 1e-12,
                  out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Ring division of the spectrum (whiten) and / or least-squares plane +
    (x - mean) / std (normalize) of [M, N, N] (or [N, N]) float32 images."""
    require_hip(images, "images")
    squeeze = images.ndim == 2
    img = images[None] if squeeze else images
    if img.ndim != 3 or img.shape[1] != img.shape[2] or img.dtype != torch.float32:
        raise ValueError("images must be float32 [M, N, N] or [N, N]")
    N, M = int(img.shape[1]), int(img.shape[0])
    if N not in SUPPORTED_BOXES:
        raise ValueError(f"image size {N} not in {SUPPORTED_BOXES}")
    if not (whiten or normalize):
        raise ValueError("nothing to do: whiten and normalize are both off")
    img = img.contiguous()
    if out is None:
        res = torch.empty_like(img)
    else:
        res = out[None] if out.ndim == 2 else out
        if tuple(res.shape) != tuple(img.shape) or res.dtype != torch.float32 or not res.is_contiguous():
            raise ValueError("out must be contiguous float32 of the images' shape")
        if res.data_ptr() == img.data_ptr():
            raise ValueError("out may not alias the input")
    wsb = int(lib().gis_build_templates_workspace(N, M))
    ws = torch.empty((max(wsb, 1),), dtype=torch.uint8, device=img.device)
    check(lib().gis_whiten_images(ptr(img), N, M, int(bool(whiten)), int(bool(normalize)), float(eps), ptr(res),
                                  ptr(ws), wsb, stream_of(img.device)), "gis_whiten_images")
    return res[0] if squeeze else res


def _spectra(dft: torch.Tensor):
    """(box N, candidates C) of one bricked spectrum ([E]) or C of them ([C, E])."""
    require_hip(dft, "dft")
    if dft.dtype != torch.complex64 or dft.ndim not in (1, 2):
        raise ValueError("dft must be a bricked complex64 spectrum [E] from volume_dft, or [C, E]")
    E = int(dft.shape[-1])
    N = next((b for b in SUPPORTED_BOXES if int(lib().gis_volume_dft_bytes(b)) == E * 8), 0)
    if not N:
        raise ValueError("dft must be a bricked spectrum from volume_dft")
    return N, (int(dft.shape[0]) if dft.ndim == 2 else 1)


def padded_row_stride(n: int) -> int:
    """Row stride (complex values) of the padded packed layout: n/2+1 rounded
    up to a multiple of 16, so every 16-column segment of a row starts on a
    128-byte line."""
    return (n // 2 + 1 + 15) // 16 * 16


def pad_rows(x: torch.Tensor, ld: int) -> torch.Tensor:
    """A contiguous copy of ``x`` [..., n/2+1] with rows ``ld`` wide, zero padded."""
    out = torch.zeros(tuple(x.shape[:-1]) + (int(ld),), dtype=x.dtype, device=x.device)
    out[..., :x.shape[-1]] = x
    return out


@dataclass
class FragmentPlanes:
    """Packed planes (n rows, n/2+1 columns) in fragment order, the searches'
    operand layout: each 16 x 16 coefficient tile one contiguous 2 KiB run
    that a wave reads with 16-byte loads.  ``data`` [..., fragment_plane(n)]
    complex64; the leading dimensions are those of the row-major planes."""
    data: torch.Tensor
    n: int

    @property
    def shape(self):
        return tuple(self.data.shape[:-1])

    def __getitem__(self, idx) -> "FragmentPlanes":
        """Planes selected along the leading dimensions (a view)."""
        return FragmentPlanes(self.data[idx], self.n)

    def unpack(self) -> torch.Tensor:
        """The planes row major, [..., n, n/2+1] (a gather; for checks)."""
        pos = torch.from_numpy(fragment_positions(self.n)).to(self.data.device)
        return self.data[..., pos.reshape(-1)].reshape(self.shape + tuple(pos.shape))


def fragment_positions(n: int) -> np.ndarray:
    """[n, n/2+1] position of each packed coefficient in fragment order."""
    m = n // 2 + 1
    nks = ((n + 15) // 16 + 1) & ~1
    lone = m % 16 == 1
    nct = m // 16 if lone else (m + 15) // 16
    r, c = np.meshgrid(np.arange(n), np.arange(m), indexing="ij")
    pos = ((r >> 4) * nct + (c >> 4)) * 256 + ((((r >> 1) & 1) * 4 + ((r >> 2) & 3)) * 16 + (c & 15)) * 2 + (r & 1)
    if lone:
        pos[:, m - 1] = nks * nct * 256 + np.arange(n)
    return pos.astype(np.int64)


def fragment_plane(n: int) -> int:
    """Complex values of one plane in fragment order at newbox n."""
    return int(lib().gis_fragment_plane(int(n)))


def pack_fragments(x: torch.Tensor, out: Optional[torch.Tensor] = None) -> FragmentPlanes:
    """``x`` [..., n, ld] (ld >= n/2+1, contiguous) in fragment order."""
    require_hip(x, "planes")
    n, ld = int(x.shape[-2]), int(x.shape[-1])
    if x.dtype != torch.complex64 or not x.is_contiguous() or ld < n // 2 + 1:
        raise ValueError("planes must be contiguous complex64 [..., n, ld >= n/2+1]")
    lead = tuple(x.shape[:-2])
    rows = int(np.prod(lead)) if lead else 1
    F = fragment_plane(n)
    if out is None:
        out = torch.empty(lead + (F,), dtype=torch.complex64, device=x.device)
    elif tuple(out.shape) != lead + (F,) or out.dtype != torch.complex64 or not out.is_contiguous():
        raise ValueError(f"out must be contiguous complex64 {lead + (F,)}")
    check(lib().gis_pack_fragments(ptr(x), rows, n, ld, ptr(out), stream_of(x.device)), "gis_pack_fragments")
    return FragmentPlanes(out, n)


H16_SCALE = 128.0  # fp16 operands: unit-norm planes times 2^7


@dataclass
class HalfFragmentPlanes:
    """Psi spectra as the fp16-operand search reads them: unit-norm planes
    times 2^7 as fp16 (real, imaginary) pairs, ``data`` [..., fragment_plane(n)]
    complex32, each tile's values one 16-byte run per lane."""
    data: torch.Tensor
    n: int

    @property
    def shape(self):
        return tuple(self.data.shape[:-1])

    def __getitem__(self, idx) -> "HalfFragmentPlanes":
        return HalfFragmentPlanes(self.data[idx], self.n)

    def unpack(self) -> torch.Tensor:
        """The planes row major as complex64 [..., n, n/2+1], still x 2^7 (for checks)."""
        pos = torch.from_numpy(half_fragment_positions(self.n)).to(self.data.device)
        return self.data.to(torch.complex64)[..., pos.reshape(-1)].reshape(self.shape + tuple(pos.shape))


def half_fragment_positions(n: int) -> np.ndarray:
    """[n, n/2+1] position of each packed coefficient in the fp16 fragment order."""
    m = n // 2 + 1
    nks = ((n + 15) // 16 + 1) & ~1
    lone = m % 16 == 1
    nct = m // 16 if lone else (m + 15) // 16
    r, c = np.meshgrid(np.arange(n), np.arange(m), indexing="ij")
    pos = ((r >> 4) * nct + (c >> 4)) * 256 + (((r >> 2) & 3) * 16 + (c & 15)) * 4 + (r & 3)
    if lone:
        pos[:, m - 1] = nks * nct * 256 + np.arange(n)
    return pos.astype(np.int64)


def pack_fragments_f16(x: FragmentPlanes, inv_norm: Optional[torch.Tensor], scale: float = H16_SCALE,
                       out: Optional[torch.Tensor] = None) -> HalfFragmentPlanes:
    """Fragment-order planes times ``scale`` * ``inv_norm`` (one value per
    plane, or None) as fp16 pairs in the fp16 fragment order."""
    if not isinstance(x, FragmentPlanes):
        raise ValueError("pack_fragments_f16 takes FragmentPlanes")
    require_hip(x.data, "planes")
    lead = x.shape
    rows = int(np.prod(lead)) if lead else 1
    F = fragment_plane(x.n)
    if inv_norm is not None:
        require_hip(inv_norm, "inv_norm")
        if inv_norm.dtype != torch.float32 or inv_norm.numel() != rows or not inv_norm.is_contiguous():
            raise ValueError("inv_norm must be contiguous float32 with one value per plane")
    if out is None:
        out = torch.empty(lead + (F,), dtype=torch.complex32, device=x.data.device)
    elif tuple(out.shape) != lead + (F,) or out.dtype != torch.complex32 or not out.is_contiguous():
        raise ValueError(f"out must be contiguous complex32 {lead + (F,)}")
    check(lib().gis_pack_fragments_f16(ptr(x.data), ptr(inv_norm), rows, x.n, float(scale), ptr(out),
                                       stream_of(x.data.device)), "gis_pack_fragments_f16")
    return HalfFragmentPlanes(out, x.n)


def build_templates(dft: torch.Tensor, new_box: int, rts: torch.Tensor, mask: Optional[torch.Tensor],
                    fsc_packed: Optional[torch.Tensor], with_norms: bool = False,
                    out: Optional[torch.Tensor] = None, normalize: bool = True,
                    workspace: Optional[torch.Tensor] = None):
    """Packed template bank [M, n, n/2+1] complex64 (unshifted), written into
    ``out`` when given ([M, n, ld] padded rows, or FragmentPlanes).  Unit-norm rows (the reference's bank) unless
    ``normalize`` is False: then the rows keep their scale and the norms are
    returned (callers fold 1/norm into the search).  ``dft`` [C, E] (C
    spectra) gives the C banks in one launch, rows c * M + d."""
    with_norms = with_norms or not normalize
    N, C = _spectra(dft)
    M = int(rts.shape[0])
    require_hip(rts, "rts")
    n = int(new_box)
    rows = C * M
    if out is None:
        bank = torch.empty((rows, n, n // 2 + 1), dtype=torch.complex64, device=dft.device)
        ldb = n // 2 + 1
    elif isinstance(out, FragmentPlanes):  # fragment order (ldb = 0)
        require_hip(out.data, "out")
        if out.n != n or tuple(out.data.shape) != (rows, fragment_plane(n)) or \
                out.data.dtype != torch.complex64 or not out.data.is_contiguous():
            raise ValueError("out must be contiguous complex64 fragment planes [C * M, fragment_plane(n)]")
        bank, ldb = out.data, 0
    else:
        require_hip(out, "out")
        if tuple(out.shape[:2]) != (rows, n) or out.ndim != 3 or int(out.shape[2]) < n // 2 + 1 \
                or out.dtype != torch.complex64 or not out.is_contiguous():
            raise ValueError("out must be contiguous complex64 [C * M, n, ld], ld >= n/2+1")
        bank = out
        ldb = int(bank.shape[2])
    norms = torch.empty((rows,), dtype=torch.float32, device=dft.device) if with_norms else None
    if mask is not None:
        require_hip(mask, "mask")
        if tuple(mask.shape) != (N, N):
            raise ValueError("mask must be [N, N]")
    if fsc_packed is not None:
        require_hip(fsc_packed, "fsc_packed")
        if tuple(fsc_packed.shape) != (n, n // 2 + 1):
            raise ValueError("fsc_packed must be [n, n/2+1]")
    ws, wsb = _template_workspace(N, rows, dft.device, workspace)
    if ldb != n // 2 + 1:  # padded rows (zeros past column n/2) or fragment order
        stride = int(dft.shape[1]) if dft.ndim > 1 else int(dft.numel())
        check(lib().gis_build_templates_batched_ld(ptr(dft), stride, C, N, n, ptr(rts), M, ptr(mask),
                                                   ptr(fsc_packed), ptr(bank), ldb, ptr(norms),
                                                   int(bool(normalize)), ptr(ws), wsb, stream_of(dft.device)),
              "gis_build_templates_batched_ld")
    elif dft.ndim == 1:
        check(lib().gis_build_templates(ptr(dft), N, n, ptr(rts), M, ptr(mask), ptr(fsc_packed), ptr(bank),
                                        ptr(norms), int(bool(normalize)), ptr(ws), wsb, stream_of(dft.device)),
              "gis_build_templates")
    else:
        check(lib().gis_build_templates_batched(ptr(dft), int(dft.shape[1]), C, N, n, ptr(rts), M, ptr(mask),
                                                ptr(fsc_packed), ptr(bank), ptr(norms), int(bool(normalize)),
                                                ptr(ws), wsb, stream_of(dft.device)), "gis_build_templates_batched")
    if ldb == 0:
        bank = out
    return (bank, norms) if with_norms else bank


def _template_workspace(N: int, rows: int, device, workspace: Optional[torch.Tensor]):
    wsb = int(lib().gis_build_templates_workspace(N, rows))
    if workspace is None:
        return torch.empty((max(wsb, 1),), dtype=torch.uint8, device=device), wsb
    require_hip(workspace, "workspace")
    if workspace.numel() * workspace.element_size() < wsb:
        raise ValueError(f"workspace must hold {wsb} bytes")
    return workspace, wsb


def build_template_coefficients(dft: torch.Tensor, new_box: int, rts: torch.Tensor, mask: Optional[torch.Tensor],
                                fsc_packed: Optional[torch.Tensor], kslot: torch.Tensor, K: int, kpad: int,
                                out: Optional[torch.Tensor] = None, norms: Optional[torch.Tensor] = None,
                                workspace: Optional[torch.Tensor] = None):
    """[M, kpad] complex64: the unit-norm bank's coefficients listed by ``kslot``
    ([n, n/2+1] int32, slot or -1), zero padded, without storing the bank;
    returns (coefficients, norms).  ``dft`` [C, E] (C spectra) gives
    [C * M, kpad] in one launch, rows c * M + d."""
    N, C = _spectra(dft)
    M = int(rts.shape[0])
    rows = C * M
    require_hip(rts, "rts")
    require_hip(kslot, "kslot")
    n = int(new_box)
    if tuple(kslot.shape) != (n, n // 2 + 1) or kslot.dtype != torch.int32:
        raise ValueError("kslot must be int32 [n, n/2+1]")
    if not 0 <= int(K) <= int(kpad):
        raise ValueError("need 0 <= K <= kpad")
    if out is None:
        out = torch.empty((rows, int(kpad)), dtype=torch.complex64, device=dft.device)
    else:
        require_hip(out, "out")
        if tuple(out.shape) != (rows, int(kpad)) or out.dtype != torch.complex64 or not out.is_contiguous():
            raise ValueError("out must be contiguous complex64 [C * M, kpad]")
    if norms is None:
        norms = torch.empty((rows,), dtype=torch.float32, device=dft.device)
    elif norms.numel() != rows or norms.dtype != torch.float32 or not norms.is_contiguous():
        raise ValueError("norms must be contiguous float32 [C * M]")
    if mask is not None:
        require_hip(mask, "mask")
        if tuple(mask.shape) != (N, N):
            raise ValueError("mask must be [N, N]")
    if fsc_packed is not None:
        require_hip(fsc_packed, "fsc_packed")
        if tuple(fsc_packed.shape) != (n, n // 2 + 1):
            raise ValueError("fsc_packed must be [n, n/2+1]")
    ws, wsb = _template_workspace(N, rows, dft.device, workspace)
    if dft.ndim == 1:
        check(lib().gis_build_template_coefficients(ptr(dft), N, n, ptr(rts), M, ptr(mask), ptr(fsc_packed),
                                                    ptr(kslot), int(K), int(kpad), ptr(out), ptr(norms), ptr(ws), wsb,
                                                    stream_of(dft.device)), "gis_build_template_coefficients")
    else:
        check(lib().gis_build_template_coefficients_batched(ptr(dft), int(dft.shape[1]), C, N, n, ptr(rts), M,
                                                            ptr(mask), ptr(fsc_packed), ptr(kslot), int(K),
                                                            int(kpad), ptr(out), ptr(norms), ptr(ws), wsb,
                                                            stream_of(dft.device)),
              "gis_build_template_coefficients_batched")
    return out, norms


def prepare_particles(images: torch.Tensor, new_box: int, rotp: torch.Tensor, mask: Optional[torch.Tensor],
                      ctf_tab: Optional[torch.Tensor], defocus: Optional[torch.Tensor], kconst: np.ndarray,
                      out: Optional[torch.Tensor] = None, inv_norm: Optional[torch.Tensor] = None):
    """Packed conjugated particle spectra [B, P, n, n/2+1] and 1/norms [B, P];
    ``out`` may be a row-major [B, P, n, ld] tensor (ld >= n/2+1) or
    FragmentPlanes [B, P] (the searches' layout, written directly)."""
    require_hip(images, "images")
    B, N = int(images.shape[0]), int(images.shape[1])
    P = int(rotp.shape[1])
    n = int(new_box)
    require_hip(rotp, "rotp")
    if tuple(rotp.shape) != (B, P, 6):
        raise ValueError("rotp must be [B, P, 6]")
    if out is None:
        out = torch.empty((B, P, n, n // 2 + 1), dtype=torch.complex64, device=images.device)
    if isinstance(out, FragmentPlanes):
        dst, ld = out.data, 0
        if out.n != n or tuple(dst.shape) != (B, P, fragment_plane(n)) or dst.dtype != torch.complex64 \
                or not dst.is_contiguous():
            raise ValueError("out must be contiguous complex64 fragment planes [B, P, fragment_plane(n)]")
    else:
        dst, ld = out, int(out.shape[-1])
        if tuple(dst.shape[:3]) != (B, P, n) or dst.dim() != 4 or ld < n // 2 + 1 or \
                dst.dtype != torch.complex64 or not dst.is_contiguous():
            raise ValueError("out must be contiguous complex64 [B, P, n, ld >= n/2+1]")
    require_hip(dst, "out")
    if inv_norm is None:
        inv_norm = torch.empty((B, P), dtype=torch.float32, device=images.device)
    kc = (C.c_float * 5)(*[float(v) for v in np.asarray(kconst, np.float32)])
    wsb = int(lib().gis_prepare_particles_workspace_ex(N, B, P))
    ws = torch.empty((max(wsb, 1),), dtype=torch.uint8, device=images.device)
    check(lib().gis_prepare_particles_ld(ptr(images), B, N, n, P, ptr(rotp), ptr(mask), ptr(ctf_tab),
                                         ptr(defocus), C.cast(kc, C.c_void_p), ptr(dst), ld, ptr(inv_norm), ptr(ws),
                                         wsb, stream_of(images.device)), "gis_prepare_particles")
    return out, inv_norm


def ctf_packed(ctf_tab: torch.Tensor, defocus: torch.Tensor, n: int, kconst: np.ndarray) -> torch.Tensor:
    """Packed weighted CTF [B, n, n/2+1] float32 of B particles."""
    require_hip(ctf_tab, "ctf_tab")
    require_hip(defocus, "defocus")
    B = int(defocus.shape[0])
    out = torch.empty((B, n, n // 2 + 1), dtype=torch.float32, device=defocus.device)
    kc = (C.c_float * 5)(*[float(v) for v in np.asarray(kconst, np.float32)])
    check(lib().gis_ctf_packed(ptr(ctf_tab), ptr(defocus), B, int(n), C.cast(kc, C.c_void_p), ptr(out),
                               stream_of(defocus.device)), "gis_ctf_packed")
    return out


@dataclass
class CorrelationResult:
    scores: torch.Tensor   # [sum M_i, P] float32
    peak_y: torch.Tensor   # [sum M_i, P] int32
    peak_x: torch.Tensor   # [sum M_i, P] int32
    best: torch.Tensor     # [B] int32 (local_dir * P + psi, -1 if none)
    peak: torch.Tensor     # [B] float32
    mean: torch.Tensor     # [B] float64
    std: torch.Tensor      # [B] float64


def _check_pmap(pmap, B, rows):
    if pmap is None:
        return
    require_hip(pmap, "pmap")
    if pmap.dtype != torch.int32 or tuple(pmap.shape) != (B,):
        raise ValueError("pmap must be int32 [B]")
    if B and (int(pmap.min()) < 0 or int(pmap.max()) >= rows):
        raise ValueError("pmap out of range of the particle rows")


def _search_operands(bank, spectra):
    """(bank, spectra, n, ldt, ldp) of the searches: row-major tensors
    [M, n, ld] / [B, P, n, ld] (ld >= n/2+1) or FragmentPlanes (ld = 0)."""
    def one(x, what, lead):
        if isinstance(x, FragmentPlanes):
            require_hip(x.data, what)
            if x.data.dim() != lead + 1 or int(x.data.shape[-1]) != fragment_plane(x.n) or \
                    not x.data.is_contiguous() or x.data.dtype != torch.complex64:
                raise ValueError(f"{what}: fragment planes must be contiguous complex64 [..., fragment_plane(n)]")
            return x.data, x.n, 0
        require_hip(x, what)
        if x.dim() != lead + 2 or not x.is_contiguous() or int(x.shape[-1]) < int(x.shape[-2]) // 2 + 1:
            raise ValueError("bank [M, n, ldt] and spectra [B, P, n, ldp] must be contiguous, ld >= n/2+1")
        return x, int(x.shape[-2]), int(x.shape[-1])
    bank, nt, ldt = one(bank, "bank", 1)
    spectra, n, ldp = one(spectra, "spectra", 2)
    if nt != n:
        raise ValueError(f"bank newbox {nt} != spectra newbox {n}")
    if (ldt == 0) != (ldp == 0):
        raise ValueError("bank and spectra row strides: both 0 (fragment order) or both row major")
    return bank, spectra, n, ldt, ldp


def correlate(bank: torch.Tensor, spectra: torch.Tensor, inv_norm: Optional[torch.Tensor],
              dir_off: torch.Tensor, dir_idx: torch.Tensor, crop: int, validate: bool = True,
              pmap: Optional[torch.Tensor] = None, tinv: Optional[torch.Tensor] = None,
              order: Optional[torch.Tensor] = None, dt: int = 0, pt: int = 0) -> CorrelationResult:
    """FP32 v606 search.  With ``pmap`` the searched particles are
    B = len(dir_off) - 1 entries each mapped to a row of ``spectra``;
    ``tinv`` gives 1/norm per bank row for an unnormalised bank; ``bank`` and
    ``spectra`` may carry zero-padded rows (``padded_row_stride``) or be
    FragmentPlanes (``pack_fragments``);
    ``order`` / ``dt`` / ``pt`` as in ``correlate_split`` (dt 1 or 2)."""
    bank, spectra, n, ldt, ldp = _search_operands(bank, spectra)
    P = int(spectra.shape[1])
    B = int(dir_off.shape[0]) - 1 if pmap is not None else int(spectra.shape[0])
    if dir_off.dtype != torch.int32 or dir_idx.dtype != torch.int32:
        raise ValueError("dir_off / dir_idx must be int32")
    if tuple(dir_off.shape) != (B + 1,):
        raise ValueError("dir_off must be [B+1]")
    total = int(dir_idx.shape[0])
    if validate:
        off = dir_off.cpu().numpy()
        if off[0] != 0 or off[-1] != total or np.any(np.diff(off) < 0):
            raise ValueError("dir_off must start at 0, be non-decreasing and end at len(dir_idx)")
        if total and (int(dir_idx.min()) < 0 or int(dir_idx.max()) >= int(bank.shape[0])):
            raise ValueError("dir_idx out of range of the template bank")
        _check_pmap(pmap, B, int(spectra.shape[0]))
        if order is not None and not torch.equal(torch.sort(order.long()).values,
                                                 torch.arange(B, device=order.device)):
            raise ValueError("order must be a permutation of the particles")
    if order is not None:
        require_hip(order, "order")
        if order.dtype != torch.int32 or tuple(order.shape) != (B,):
            raise ValueError("order must be int32 [B]")
    if tinv is not None:
        require_hip(tinv, "tinv")
        if tinv.dtype != torch.float32 or tinv.numel() != int(bank.shape[0]):
            raise ValueError("tinv must be float32 with one value per bank row")
    dev = spectra.device
    scores = torch.empty((total, P), dtype=torch.float32, device=dev)
    py = torch.empty((total, P), dtype=torch.int32, device=dev)
    px = torch.empty((total, P), dtype=torch.int32, device=dev)
    best = torch.empty((B,), dtype=torch.int32, device=dev)
    peak = torch.empty((B,), dtype=torch.float32, device=dev)
    mean = torch.empty((B,), dtype=torch.float64, device=dev)
    std = torch.empty((B,), dtype=torch.float64, device=dev)
    wsb = int(lib().gis_correlate_workspace(n, int(crop), B))
    ws = torch.empty((max(wsb, 1),), dtype=torch.uint8, device=dev)
    check(lib().gis_correlate(ptr(bank), ptr(spectra), ptr(inv_norm), ptr(tinv), ptr(dir_off), ptr(dir_idx),
                              ptr(pmap), ptr(order), B, P, n, int(crop), total, ldt, ldp, int(dt), int(pt),
                              ptr(scores), ptr(py), ptr(px), ptr(best), ptr(peak), ptr(mean), ptr(std), ptr(ws),
                              wsb, stream_of(dev)), "gis_correlate")
    return CorrelationResult(scores, py, px, best, peak, mean, std)


def correlate_split(bank: torch.Tensor, spectra: torch.Tensor, inv_norm: Optional[torch.Tensor],
                    dir_off: torch.Tensor, dir_idx: torch.Tensor, crop: int, validate: bool = True,
                    pmap: Optional[torch.Tensor] = None, tinv: Optional[torch.Tensor] = None,
                    order: Optional[torch.Tensor] = None, dt: int = 0, pt: int = 0) -> CorrelationResult:
    """The v606 search of ``correlate`` (every packed coefficient, every lag of
    the crop, crop <= 32) on split-bf16 MFMA.  ``bank`` [M, n, ldt] and
    ``spectra`` [B, P, n, ldp] may carry zero-padded rows (ld > n/2+1;
    ``padded_row_stride(n)`` aligns them to 128-byte lines) or be
    FragmentPlanes (``pack_fragments``: 16-byte loads).  ``order`` is the
    particles' processing order (a permutation of the B searched particles, e.g. by
    orientation so neighbouring particles share template rows in cache);
    ``dt`` / ``pt`` are the directions / psi values per wave (0 = automatic;
    results do not depend on them)."""
    bank, spectra, n, ldt, ldp = _search_operands(bank, spectra)
    P = int(spectra.shape[1])
    B = int(dir_off.shape[0]) - 1 if pmap is not None else int(spectra.shape[0])
    if dir_off.dtype != torch.int32 or dir_idx.dtype != torch.int32:
        raise ValueError("dir_off / dir_idx must be int32")
    if tuple(dir_off.shape) != (B + 1,):
        raise ValueError("dir_off must be [B+1]")
    total = int(dir_idx.shape[0])
    if order is not None:
        require_hip(order, "order")
        if order.dtype != torch.int32 or tuple(order.shape) != (B,):
            raise ValueError("order must be int32 [B]")
    if validate:
        off = dir_off.cpu().numpy()
        if off[0] != 0 or off[-1] != total or np.any(np.diff(off) < 0):
            raise ValueError("dir_off must start at 0, be non-decreasing and end at len(dir_idx)")
        if total and (int(dir_idx.min()) < 0 or int(dir_idx.max()) >= int(bank.shape[0])):
            raise ValueError("dir_idx out of range of the template bank")
        _check_pmap(pmap, B, int(spectra.shape[0]))
        if order is not None and not torch.equal(torch.sort(order.long()).values,
                                                 torch.arange(B, device=order.device)):
            raise ValueError("order must be a permutation of the particles")
    if tinv is not None:
        require_hip(tinv, "tinv")
        if tinv.dtype != torch.float32 or tinv.numel() != int(bank.shape[0]):
            raise ValueError("tinv must be float32 with one value per bank row")
    dev = spectra.device
    scores = torch.empty((total, P), dtype=torch.float32, device=dev)
    py = torch.empty((total, P), dtype=torch.int32, device=dev)
    px = torch.empty((total, P), dtype=torch.int32, device=dev)
    best = torch.empty((B,), dtype=torch.int32, device=dev)
    peak = torch.empty((B,), dtype=torch.float32, device=dev)
    mean = torch.empty((B,), dtype=torch.float64, device=dev)
    std = torch.empty((B,), dtype=torch.float64, device=dev)
    wsb = int(lib().gis_correlate_split_workspace(n, int(crop), B))
    ws = torch.empty((max(wsb, 1),), dtype=torch.uint8, device=dev)
    check(lib().gis_correlate_split(ptr(bank), ptr(spectra), ptr(inv_norm), ptr(tinv), ptr(dir_off), ptr(dir_idx),
                                    ptr(pmap), ptr(order), B, P, n, int(crop), total, ldt, ldp, int(dt), int(pt),
                                    ptr(scores), ptr(py), ptr(px), ptr(best), ptr(peak), ptr(mean), ptr(std),
                                    ptr(ws), wsb, stream_of(dev)), "gis_correlate_split")
    return CorrelationResult(scores, py, px, best, peak, mean, std)


def correlate_split_f16(bank: FragmentPlanes, spectra: HalfFragmentPlanes, dir_off: torch.Tensor,
                        dir_idx: torch.Tensor, crop: int, validate: bool = True, pmap: Optional[torch.Tensor] = None,
                        tinv: Optional[torch.Tensor] = None, order: Optional[torch.Tensor] = None, dt: int = 0,
                        pt: int = 0, wsplit: bool = False, refine: Optional[FragmentPlanes] = None,
                        inv_norm: Optional[torch.Tensor] = None,
                        bank16: Optional[HalfFragmentPlanes] = None,
                        out: Optional[CorrelationResult] = None) -> CorrelationResult:
    """The v606 search of ``correlate`` with fp16 operands: ``bank``
    float32 fragment planes [M] (``tinv`` 1/norm per row, or None for unit
    rows), ``spectra`` from ``pack_fragments_f16`` (unit norm x 2^7),
    products in packed fp16, the row table split into fp16 hi + lo
    (``wsplit``) or fp16 alone; ``order`` / ``dt`` / ``pt`` as in
    ``correlate_split``.  ``refine``: the float32 fragment-order spectra
    (with ``inv_norm``) the fp16 ones came from; each particle's best pair
    and its near ties are then rescored on split-bf16 MFMA.  ``bank16``: the
    bank as fp16 pairs (``pack_fragments_f16(bank, tinv)``), read by the fp16
    pass instead of rounding the float32 rows per step.  ``out``: result
    tensors to write into (shapes and dtypes as returned)."""
    if not isinstance(bank, FragmentPlanes) or not isinstance(spectra, HalfFragmentPlanes):
        raise ValueError("correlate_split_f16 takes FragmentPlanes (bank) and HalfFragmentPlanes (spectra)")
    require_hip(bank.data, "bank")
    require_hip(spectra.data, "spectra")
    n = spectra.n
    if bank.n != n:
        raise ValueError(f"bank newbox {bank.n} != spectra newbox {n}")
    F = fragment_plane(n)
    if bank.data.dim() != 2 or int(bank.data.shape[-1]) != F or bank.data.dtype != torch.complex64 or \
            not bank.data.is_contiguous():
        raise ValueError("bank must be contiguous complex64 [M, fragment_plane(n)]")
    if spectra.data.dim() != 3 or int(spectra.data.shape[-1]) != F or spectra.data.dtype != torch.complex32 or \
            not spectra.data.is_contiguous():
        raise ValueError("spectra must be contiguous complex32 [B, P, fragment_plane(n)]")
    P = int(spectra.data.shape[1])
    B = int(dir_off.shape[0]) - 1 if pmap is not None else int(spectra.data.shape[0])
    if dir_off.dtype != torch.int32 or dir_idx.dtype != torch.int32:
        raise ValueError("dir_off / dir_idx must be int32")
    if tuple(dir_off.shape) != (B + 1,):
        raise ValueError("dir_off must be [B+1]")
    total = int(dir_idx.shape[0])
    if order is not None:
        require_hip(order, "order")
        if order.dtype != torch.int32 or tuple(order.shape) != (B,):
            raise ValueError("order must be int32 [B]")
    if validate:
        off = dir_off.cpu().numpy()
        if off[0] != 0 or off[-1] != total or np.any(np.diff(off) < 0):
            raise ValueError("dir_off must start at 0, be non-decreasing and end at len(dir_idx)")
        if total and (int(dir_idx.min()) < 0 or int(dir_idx.max()) >= int(bank.data.shape[0])):
            raise ValueError("dir_idx out of range of the template bank")
        _check_pmap(pmap, B, int(spectra.data.shape[0]))
        if order is not None and not torch.equal(torch.sort(order.long()).values,
                                                 torch.arange(B, device=order.device)):
            raise ValueError("order must be a permutation of the particles")
    if tinv is not None:
        require_hip(tinv, "tinv")
        if tinv.dtype != torch.float32 or tinv.numel() != int(bank.data.shape[0]):
            raise ValueError("tinv must be float32 with one value per bank row")
    dev = spectra.data.device
    if out is None:
        out = CorrelationResult(torch.empty((total, P), dtype=torch.float32, device=dev),
                                torch.empty((total, P), dtype=torch.int32, device=dev),
                                torch.empty((total, P), dtype=torch.int32, device=dev),
                                torch.empty((B,), dtype=torch.int32, device=dev),
                                torch.empty((B,), dtype=torch.float32, device=dev),
                                torch.empty((B,), dtype=torch.float64, device=dev),
                                torch.empty((B,), dtype=torch.float64, device=dev))
    want = [((total, P), torch.float32), ((total, P), torch.int32), ((total, P), torch.int32), ((B,), torch.int32),
            ((B,), torch.float32), ((B,), torch.float64), ((B,), torch.float64)]
    fields = [out.scores, out.peak_y, out.peak_x, out.best, out.peak, out.mean, out.std]
    for t, (shape, dtype) in zip(fields, want):
        if tuple(t.shape) != shape or t.dtype != dtype or not t.is_contiguous() or t.device != dev:
            raise ValueError("out tensors must match the result's shapes and dtypes")
    scores, py, px, best, peak, mean, std = fields
    b16 = None
    if bank16 is not None:
        if not isinstance(bank16, HalfFragmentPlanes) or bank16.n != n or \
                tuple(bank16.data.shape) != tuple(bank.data.shape) or bank16.data.dtype != torch.complex32:
            raise ValueError("bank16 must be the bank packed by pack_fragments_f16")
        require_hip(bank16.data, "bank16")
        b16 = bank16.data
    s32 = None
    if refine is not None:
        if not isinstance(refine, FragmentPlanes) or refine.n != n or \
                tuple(refine.data.shape) != tuple(spectra.data.shape) or refine.data.dtype != torch.complex64:
            raise ValueError("refine must be the float32 FragmentPlanes the fp16 spectra were packed from")
        if inv_norm is None:
            raise ValueError("refine needs the spectra's inv_norm")
        require_hip(refine.data, "refine")
        s32 = refine.data
    wsb = int(lib().gis_correlate_split_f16_workspace(n, int(crop), B, total * P))
    ws = torch.empty((max(wsb, 1),), dtype=torch.uint8, device=dev)
    check(lib().gis_correlate_split_f16(ptr(bank.data), ptr(b16), ptr(spectra.data), ptr(s32), ptr(inv_norm), ptr(tinv),
                                        ptr(dir_off), ptr(dir_idx), ptr(pmap), ptr(order), B, P, n, int(crop), total,
                                        int(dt), int(pt), 1 if wsplit else 0, ptr(scores), ptr(py), ptr(px),
                                        ptr(best), ptr(peak), ptr(mean), ptr(std), ptr(ws), wsb, stream_of(dev)),
          "gis_correlate_split_f16")
    return CorrelationResult(scores, py, px, best, peak, mean, std)


def score_sum(peak: torch.Tensor, best: Optional[torch.Tensor] = None) -> torch.Tensor:
    """sum exp(peak/1024) over particles with best >= 0, as a float64 device scalar."""
    require_hip(peak, "peak")
    out = torch.empty((1,), dtype=torch.float64, device=peak.device)
    check(lib().gis_score_sum(ptr(peak), ptr(best), int(peak.shape[0]), ptr(out), stream_of(peak.device)),
          "gis_score_sum")
    return out


def score_sum_segments(peak: torch.Tensor, best: Optional[torch.Tensor], off: torch.Tensor) -> torch.Tensor:
    """Per-segment sum exp(peak/1024) (segments [off[s], off[s+1])) as float64 [nseg]."""
    require_hip(peak, "peak")
    require_hip(off, "off")
    if off.dtype != torch.int32:
        raise ValueError("off must be int32")
    nseg = int(off.shape[0]) - 1
    out = torch.empty((max(nseg, 0),), dtype=torch.float64, device=peak.device)
    check(lib().gis_score_sum_segments(ptr(peak), ptr(best), ptr(off), nseg, ptr(out), stream_of(peak.device)),
          "gis_score_sum_segments")
    return out


def gather_coefficients(src: torch.Tensor, kidx: torch.Tensor, kpad: int,
                        scale: Optional[torch.Tensor] = None, out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """[rows, Kpad] complex64: the listed packed coefficients of each spectrum of
    ``src`` ([..., n, n/2+1]), times ``scale`` per row if given, zero padded."""
    require_hip(src, "src")
    require_hip(kidx, "kidx")
    if src.dtype != torch.complex64 or kidx.dtype != torch.int32:
        raise ValueError("src must be complex64 and kidx int32")
    plane = int(src.shape[-1]) * int(src.shape[-2])
    rows = src.numel() // plane if plane else 0
    K = int(kidx.numel())
    if kpad < K or kpad % 16:
        raise ValueError("kpad must be a multiple of 16 and >= len(kidx)")
    if scale is not None:
        require_hip(scale, "scale")
        if scale.dtype != torch.float32 or scale.numel() != rows:
            raise ValueError("scale must be float32 with one value per spectrum")
    if out is None:
        out = torch.empty((rows, kpad), dtype=torch.complex64, device=src.device)
    else:
        require_hip(out, "out")
        if tuple(out.shape) != (rows, kpad) or out.dtype != torch.complex64:
            raise ValueError("out must be complex64 [rows, kpad]")
    check(lib().gis_gather_coefficients(ptr(src), plane, rows, ptr(kidx), K, int(kpad), ptr(scale), ptr(out),
                                        stream_of(src.device)), "gis_gather_coefficients")
    return out


# Largest |T_k P_k| the fp16 operand represents after its 2^14 scale (65504 / 2^14).
FP16_PRODUCT_BOUND = 65504.0 / 16384.0


def correlate_mixed(tk: torch.Tensor, pk: torch.Tensor, P: int, dir_off: torch.Tensor, dir_idx: torch.Tensor,
                    b_hi: torch.Tensor, b_lo: torch.Tensor, half: int, validate: bool = True,
                    pmap: Optional[torch.Tensor] = None, pt: int = 0, wpb: int = 0,
                    plast: Optional[torch.Tensor] = None, rows: Optional[torch.Tensor] = None,
                    operand: str = "bf16") -> CorrelationResult:
    """Split-bf16 (``operand="fp16"``: split-fp16) small-shift search: best
    lag over [-half, half]^2 of every (direction, psi) pair; peak_y / peak_x
    are lag + half.  ``b_hi`` / ``b_lo`` are the lag table split in that
    operand type (``host.SmallShiftTables(..., operand=...)``).  ``validate``
    checks dir_off / dir_idx against the operands on the host, and for fp16
    that |T_k P_k| <= 65504 / 2^14; callers with ``validate=False`` must pass
    unit-norm coefficient rows (every |T_k|, |P_k| <= 1), which is what
    ``gather_coefficients`` with 1/norm scales and the template-coefficient
    kernel produce."""
    if operand not in ("bf16", "fp16"):
        raise ValueError(f"operand {operand!r} not 'bf16' or 'fp16'")
    for t, name in ((tk, "tk"), (pk, "pk"), (dir_off, "dir_off"), (dir_idx, "dir_idx"), (b_hi, "b_hi"),
                    (b_lo, "b_lo")):
        require_hip(t, name)
    kpad = int(tk.shape[1])
    if int(pk.shape[1]) != kpad or pk.dtype != torch.complex64 or tk.dtype != torch.complex64:
        raise ValueError("tk and pk must be complex64 with the same Kpad")
    B = int(dir_off.shape[0]) - 1
    if pmap is None and int(pk.shape[0]) != B * P:
        raise ValueError("pk must hold P rows per particle")
    if int(pk.shape[0]) % P:
        raise ValueError("pk must hold P rows per particle")
    if tuple(b_hi.shape) != tuple(b_lo.shape) or b_hi.numel() != 4 * (kpad // 16) * 64 * 8:
        raise ValueError("lag table does not match Kpad")
    nlag = (2 * int(half) + 1) ** 2
    if nlag % 16 == 1 and nlag > 1:
        if plast is None:
            raise ValueError(f"{nlag} lags need the last-lag table (plast)")
        require_hip(plast, "plast")
        if plast.dtype != torch.float32 or plast.numel() != 2 * kpad:
            raise ValueError("plast must be float32 [Kpad, 2]")
    if dir_off.dtype != torch.int32 or dir_idx.dtype != torch.int32:
        raise ValueError("dir_off / dir_idx must be int32")
    total = int(dir_idx.shape[0])
    if validate:
        off = dir_off.cpu().numpy().astype(np.int64)
        if off[0] != 0 or off[-1] != total or np.any(np.diff(off) < 0):
            raise ValueError("dir_off must start at 0, be non-decreasing and end at len(dir_idx)")
        if total and (int(dir_idx.min()) < 0 or int(dir_idx.max()) >= int(tk.shape[0])):
            raise ValueError("dir_idx out of range of the template coefficients")
        _check_pmap(pmap, B, int(pk.shape[0]) // P)
        if rows is not None and not torch.equal(torch.sort(rows.long()).values,
                                                torch.arange(total, device=rows.device)):
            raise ValueError("rows must be a permutation of the dir_idx rows")
        # fp16 holds 2^14 |T_k P_k| only while it stays at most 65504
        if operand == "fp16" and tk.numel() and pk.numel() and \
                float(tk.abs().max()) * float(pk.abs().max()) > FP16_PRODUCT_BOUND:
            raise ValueError(f"the fp16 operand needs |T_k P_k| <= {FP16_PRODUCT_BOUND} (unit-norm spectra)")
    if rows is not None:
        require_hip(rows, "rows")
        if rows.dtype != torch.int32 or rows.numel() != total:
            raise ValueError("rows must be int32 [len(dir_idx)]")
    dev = pk.device
    scores = torch.empty((total, P), dtype=torch.float32, device=dev)
    py = torch.empty((total, P), dtype=torch.int32, device=dev)
    px = torch.empty((total, P), dtype=torch.int32, device=dev)
    best = torch.empty((B,), dtype=torch.int32, device=dev)
    peak = torch.empty((B,), dtype=torch.float32, device=dev)
    mean = torch.empty((B,), dtype=torch.float64, device=dev)
    std = torch.empty((B,), dtype=torch.float64, device=dev)
    check(lib().gis_correlate_mixed_ex(ptr(tk), ptr(pk), ptr(dir_off), ptr(dir_idx), ptr(pmap), ptr(rows), B,
                                       int(P), total, kpad, ptr(b_hi), ptr(b_lo), ptr(plast), int(half), int(pt),
                                       int(wpb), ptr(scores), ptr(py), ptr(px), ptr(best), ptr(peak), ptr(mean),
                                       ptr(std), 0 if operand == "bf16" else 1, stream_of(dev)),
          "gis_correlate_mixed")
    return CorrelationResult(scores, py, px, best, peak, mean, std)
