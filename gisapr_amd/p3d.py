"""Standalone projector: real-space central-slice projections of a model
volume at the orientations of a RELION STAR file, optionally whitened
(project3d_and_whiten_cuda.py, same CLI and library names).

Every projection comes from the template kernel's real-space modes
(``kernels.project_slices``): one launch for all orientations, with the pad /
sinc^2 / rfftn prefix of ``kernels.volume_dft`` (unpadded for
``--skip_padding``).  ``getSpectrum_divideBySpectrum_torch`` and
``normalize_torch`` run the same ring division and plane normalisation on
given images (``kernels.whiten_images``); ``extract_central_slices_rfft``
samples central slices of a natural-layout spectrum (``kernels.extract_slices``).
``run_projection`` and ``apply_sinc2_filter_inplace_3d`` are the library form
of project3d_and_whiten_cuda_nowrite.py.
There is no CPU path: tensors must live on the HIP device.
"""
from __future__ import annotations

import argparse
import os
import time
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import host, kernels, mrc, star
from .projection import apply_sinc2_filter_inplace_3d  # noqa: F401  (project3d_and_whiten_cuda_nowrite.py:42-81)

__all__ = ["create_project3d_parser", "Euler_angles2matrix", "DEG2RAD", "build_radius_map_torch",
           "build_xy_grid_torch", "getSpectrum_divideBySpectrum_torch", "normalize_torch",
           "extract_central_slices_rfft", "apply_sinc2_filter_inplace_3d", "run_project3d", "run_projection",
           "extract_filename", "main", "judge_relion30_or_relion31", "judge_mline0", "judge_mline1",
           "rfft_shape", "dft_center", "fftshift_2d", "fftshift_3d", "fftfreq_grid", "central_slice_grid",
           "rotated_central_slice_grid", "fftfreq_to_dft_coordinates", "array_to_grid_sample", "sample_dft_3d"]


def create_project3d_parser() -> argparse.ArgumentParser:
    """Flags of project3d_and_whiten_cuda.py:15-23 and the --return_array
    flag of project3d_and_whiten_cuda_nowrite.py:25."""
    parser = argparse.ArgumentParser(description="Project models from relion angle star then whitening them.")
    parser.add_argument("--i", type=str, required=True, help="Input model in mrc form.")
    parser.add_argument("--ang", type=str, required=True, help="The RELION angle starfile.")
    parser.add_argument("--o", type=str, required=True, help="The output starfile name.")
    parser.add_argument("--skip_padding", action="store_true", help="Skip the padding process. default = False")
    parser.add_argument("--normal_background_powerspectrum", action="store_true",
                        help="Apply whitening filter to output projections. default = False")
    parser.add_argument("--gpuid", type=str, default=None,
                        help="The gpuid, only one gpuid should be given. If not given, the current HIP device.")
    parser.add_argument("--return_array", action="store_true",
                        help="Return numpy array instead of writing the stack. default = False")
    return parser


def Euler_angles2matrix(alpha, beta, gamma) -> List[List[float]]:
    """RELION ZYZ Euler matrix in degrees with pi written 3.14159265359
    (project3d_and_whiten_cuda.py:700-722)."""
    return host.euler_angles_to_matrix(float(alpha), float(beta), float(gamma))


def build_radius_map_torch(ysize: int, xsize: int, device) -> torch.Tensor:
    """round(sqrt((y - ysize//2)^2 + (x - xsize//2)^2)) as int64 (:24-31)."""
    y = torch.arange(ysize, device=device, dtype=torch.float32)
    x = torch.arange(xsize, device=device, dtype=torch.float32)
    yy, xx = torch.meshgrid(y, x, indexing="ij")
    return torch.round(torch.sqrt((yy - ysize // 2) ** 2 + (xx - xsize // 2) ** 2)).long()


def build_xy_grid_torch(ysize: int, xsize: int, device):
    """(row index grid, column index grid) in float32 (:33-37)."""
    y = torch.arange(ysize, device=device, dtype=torch.float32)
    x = torch.arange(xsize, device=device, dtype=torch.float32)
    yy, xx = torch.meshgrid(y, x, indexing="ij")
    return yy, xx


def _image_2d(img: torch.Tensor, name: str) -> torch.Tensor:
    if not isinstance(img, torch.Tensor) or img.ndim != 2:
        raise ValueError(f"{name} must be 2D")
    if img.shape[0] != img.shape[1]:
        raise ValueError(f"{name}: only square images are supported (got {tuple(img.shape)})")
    kernels.require_hip(img, name)
    return img.to(torch.float32).contiguous()


def getSpectrum_divideBySpectrum_torch(img: torch.Tensor, radius_map: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Ring-mean |F| whitening of one real image (:196-230): divide the
    centred spectrum by its ring mean (DC ring untouched), inverse, real part.
    ``radius_map``, when given, must be the default ring map."""
    im = _image_2d(img, "img")
    if radius_map is not None:
        want = build_radius_map_torch(im.shape[0], im.shape[1], im.device)
        if tuple(radius_map.shape) != tuple(want.shape) or not torch.equal(radius_map.to(want.device).long(), want):
            raise ValueError("only the default ring map round(sqrt((y-H//2)^2 + (x-W//2)^2)) is supported")
    return kernels.whiten_images(im, whiten=True, normalize=False)


def normalize_torch(img: torch.Tensor, xx: Optional[torch.Tensor] = None, yy: Optional[torch.Tensor] = None,
                    eps: float = 1e-12) -> torch.Tensor:
    """Least-squares plane removal then (x - mean) / std, or x - mean when
    std <= eps (:231-291).  ``xx`` / ``yy``, when given, must be the index
    grids of the image (in either role: a plane fitted over the whole grid
    does not depend on which grid is called x)."""
    im = _image_2d(img, "img")
    if (xx is None) != (yy is None):
        raise ValueError("give both xx and yy, or neither")
    if xx is not None:
        gy, gx = build_xy_grid_torch(im.shape[0], im.shape[1], im.device)
        a, b = xx.to(im.device, torch.float32), yy.to(im.device, torch.float32)
        if not ((torch.equal(a, gx) and torch.equal(b, gy)) or (torch.equal(a, gy) and torch.equal(b, gx))):
            raise ValueError("xx / yy must be the image's row and column index grids")
    return kernels.whiten_images(im, whiten=False, normalize=True, eps=eps)


def extract_central_slices_rfft(dft: torch.Tensor, image_shape, rotation_matrices: torch.Tensor,
                                rotation_matrix_zyx: bool) -> torch.Tensor:
    """Central slices of an fftshifted rfft (:397-435): ``dft`` [L, L, L/2+1]
    with planes and rows fftshifted, ``image_shape`` the real (L, L, L)
    shape, ``rotation_matrices`` [..., 3, 3] acting on xyz vectors (or zyx
    when ``rotation_matrix_zyx``).  Returns [..., L, L/2+1] complex64 with
    rows fftshifted."""
    kernels.require_hip(dft, "dft")
    L = int(dft.shape[0])
    if tuple(int(v) for v in image_shape) != (L, L, L) or tuple(dft.shape) != (L, L, L // 2 + 1):
        raise ValueError("dft must be the [L, L, L/2+1] rfft of a cubic image of shape (L, L, L)")
    R = torch.as_tensor(rotation_matrices).to(dft.device, torch.float32)
    lead = tuple(R.shape[:-2])
    R = R.reshape(-1, 3, 3)
    if rotation_matrix_zyx:
        R = R.flip(-1).flip(-2)  # the same rotation acting on xyz vectors
    natural = torch.fft.ifftshift(dft.to(torch.complex64), dim=(0, 1))
    out = kernels.extract_slices(kernels.natural_to_bricked(natural.contiguous()), L, R.contiguous())
    return out.reshape(*lead, L, L // 2 + 1)


# ---------------------------------------------------------------------------
# Frequency-grid geometry of the projector script (the libtilt helpers of
# project3d_and_whiten_cuda.py:292-650).  Small index tables and coordinate
# maps on the caller's device; the sampling itself is a HIP kernel.
# ---------------------------------------------------------------------------

def DEG2RAD(x):
    """Degrees to radians with pi written 3.14159265359 (:692-693)."""
    return x / 180.0 * 3.14159265359


def rfft_shape(input_shape: Sequence[int]) -> tuple:
    """Shape after an rfft over the last axis: last = int(last / 2 + 1) (:298-302)."""
    out = list(input_shape)
    out[-1] = int(out[-1] / 2 + 1)
    return tuple(out)


def fftshift_2d(input: torch.Tensor, rfft: bool) -> torch.Tensor:  # noqa: A002 (reference name)
    """fftshift over the last two axes, or only the rows of an rfft (:292-297)."""
    return torch.fft.fftshift(input, dim=(-2,) if rfft else (-2, -1))


def fftshift_3d(input: torch.Tensor, rfft: bool) -> torch.Tensor:  # noqa: A002 (reference name)
    """fftshift over the last three axes, or all but the last of an rfft (:444-449)."""
    return torch.fft.fftshift(input, dim=(-3, -2) if rfft else (-3, -2, -1))


def _axis_freqs(n: int, spacing: float, half: bool, device) -> torch.Tensor:
    return (torch.fft.rfftfreq if half else torch.fft.fftfreq)(n, d=spacing, device=device)


def fftfreq_grid(image_shape, rfft: bool, fftshift: bool = False, spacing=1, norm: bool = False,
                 device=None) -> torch.Tensor:
    """[*shape, ndim] DFT sample frequencies of a 2-D or 3-D image (the last
    axis rfftfreq when ``rfft``), optionally fftshifted, or their Euclidean
    norm [*shape] when ``norm`` (:303-372)."""
    shape = tuple(int(v) for v in image_shape)
    if len(shape) not in (2, 3):
        raise NotImplementedError("Construction of fftfreq grids is currently only supported for 2D and 3D images.")
    sp = list(spacing) if isinstance(spacing, Sequence) else [spacing] * len(shape)
    axes = [_axis_freqs(n, float(d), rfft and i == len(shape) - 1, device) for i, (n, d) in enumerate(zip(shape, sp))]
    grid = torch.stack(torch.meshgrid(*axes, indexing="ij"), dim=-1)
    if fftshift:
        grid = torch.movedim((fftshift_2d if len(shape) == 2 else fftshift_3d)(torch.movedim(grid, -1, 0), rfft), 0,
                             -1)
    if norm:
        grid = (grid ** 2).sum(dim=-1) ** 0.5
    return grid


def central_slice_grid(image_shape, rfft: bool, fftshift: bool = False, device=None) -> torch.Tensor:
    """[h, w(/2+1), 3] (z = 0, y, x) frequencies of the central xy slice (:450-468)."""
    h, w = (int(v) for v in tuple(image_shape)[-2:])
    hw = fftfreq_grid((h, w), rfft=rfft, fftshift=fftshift, device=device)
    return torch.cat([torch.zeros_like(hw[..., :1]), hw], dim=-1)


def rotated_central_slice_grid(image_shape, rotation_matrices: torch.Tensor, rotation_matrix_zyx: bool, rfft: bool,
                               fftshift: bool = False, device=None) -> torch.Tensor:
    """[..., h, w, 3] (z, y, x) central-slice frequencies rotated by each
    matrix (acting on zyx vectors, or on xyz when not ``rotation_matrix_zyx``)
    (:469-489)."""
    grid = central_slice_grid(image_shape, rfft=rfft, fftshift=fftshift, device=device)
    R = torch.as_tensor(rotation_matrices).to(grid.device, grid.dtype)
    if not rotation_matrix_zyx:
        R = R.flip(-1).flip(-2)  # the same rotation acting on zyx vectors
    return torch.einsum("...ij,hwj->...hwi", R, grid)


def dft_center(image_shape, rfft: bool, fftshifted: bool, device=None) -> torch.Tensor:
    """Array index of the DFT origin (:524-537): shape // 2 when fftshifted
    (of the rfft shape when ``rfft``), last axis 0 for an rfft."""
    shape = rfft_shape(image_shape) if rfft else tuple(image_shape)
    c = [int(v) // 2 if fftshifted else 0 for v in shape]
    if rfft:
        c[-1] = 0
    return torch.tensor(c, dtype=torch.long, device=device)


def fftfreq_to_dft_coordinates(frequencies: torch.Tensor, image_shape, rfft: bool) -> torch.Tensor:
    """Frequencies [..., d] to array coordinates in a fftshifted DFT (:538-571):
    f * size (the last axis f * 2 (w_rfft - 1) for an rfft) plus the origin."""
    shape = [float(v) for v in image_shape]
    scale = torch.tensor(shape, dtype=frequencies.dtype, device=frequencies.device)
    if rfft:
        scale[-1] = 2.0 * (rfft_shape(image_shape)[-1] - 1)
    dc = dft_center(image_shape, rfft=rfft, fftshifted=True, device=frequencies.device)
    return frequencies * scale + dc


def array_to_grid_sample(array_coordinates: torch.Tensor, array_shape) -> torch.Tensor:
    """Array coordinates [..., d] to grid_sample's align_corners=True grid,
    c / (size/2 - 1/2) - 1, in reversed (x, y, z) order (:573-599)."""
    shape = torch.as_tensor(array_shape, dtype=array_coordinates.dtype, device=array_coordinates.device)
    return torch.flip(array_coordinates / (0.5 * shape - 0.5) - 1, dims=(-1,))


def sample_dft_3d(dft: torch.Tensor, coordinates: torch.Tensor) -> torch.Tensor:
    """Trilinear, border-clamped samples [...] of a complex [D, H, W] volume at
    array coordinates [..., 3] (z, y, x) (:612-650), one HIP launch."""
    return kernels.sample_dft_3d(dft.to(torch.complex64).contiguous(), coordinates)


# ---------------------------------------------------------------------------
# STAR helpers of the projector script (:651-691)
# ---------------------------------------------------------------------------

def judge_relion30_or_relion31(inline: Sequence[str]) -> int:
    """1 for a RELION 3.0 STAR, 0 for 3.1 (:651-665): among the first three
    lines, a '# version N' line with N > 30000 marks 3.1; a '# ... 3.0.8'
    header ends the check as 3.0."""
    for line in list(inline)[:3]:
        tok = line.split()
        if len(tok) < 2:
            continue
        if tok[0] == "#" and len(tok) >= 4 and tok[3] == "3.0.8":
            break
        if int(tok[2]) > 30000:
            return 0
    return 1


def _first_data_after_labels(inline: Sequence[str], start: int, stop: int):
    seen_label = False
    for i in range(start, min(stop, len(inline))):
        tok = inline[i].split()
        if not tok:
            continue
        if tok[0][0] == "_":
            seen_label = True
        elif seen_label:
            return i
    return None


def judge_mline0(inline: Sequence[str]):
    """Index of the first data line after the first run of '_' label lines,
    searched in the first 60 lines (:666-678); None when there is none."""
    return _first_data_after_labels(inline, 0, 60)


def judge_mline1(inline: Sequence[str], start: int):
    """judge_mline0 from line ``start``, searched up to line 70 (:679-691)."""
    return _first_data_after_labels(inline, int(start), 70)


def extract_filename(file_path: str) -> str:
    """Base name without its last extension (:724-729)."""
    base = os.path.basename(file_path)
    parts = base.split(".")
    return ".".join(parts[:-1]) if len(parts) > 1 else base


def _device(gpuid) -> torch.device:
    if gpuid is not None:
        dev = torch.device("cuda", int(gpuid))
        torch.cuda.set_device(dev)
        return dev
    if not torch.cuda.is_available():
        raise RuntimeError("the projector runs on the HIP device only; no GPU is visible")
    return torch.device("cuda", torch.cuda.current_device())


def _read_angles(particle_star: str):
    lines = star.read_text_lines(particle_star)
    loop = star.find_star_loop(lines, ("_rlnImageName", "_rlnAngleRot", "_rlnAngleTilt", "_rlnAnglePsi"))
    c_rot, c_tilt = loop.labels["_rlnAngleRot"], loop.labels["_rlnAngleTilt"]
    rows = loop.data_line_indices
    rot = np.array([float(loop.tokens(r)[c_rot]) for r in rows])
    tilt = np.array([float(loop.tokens(r)[c_tilt]) for r in rows])
    return lines, loop, rot, tilt


def _load_volume(mrc_volume) -> np.ndarray:
    vol = mrc_volume if isinstance(mrc_volume, np.ndarray) else mrc.read(mrc_volume)
    vol = np.ascontiguousarray(vol, np.float32)
    if vol.ndim != 3 or len(set(vol.shape)) != 1:
        raise ValueError(f"the model must be a cubic volume (got {vol.shape})")
    return vol


def _project(vol: np.ndarray, rot, tilt, pad: bool, whiten: bool, dev: torch.device) -> np.ndarray:
    """[M, N, N] float32 projections: one volume transform, one template
    kernel launch for every orientation (whitened in the same launch)."""
    if len(rot) == 0:
        return np.zeros((0,) + vol.shape[1:], np.float32)
    dft = kernels.volume_dft(torch.from_numpy(vol).to(dev), pad=pad)
    rts = torch.from_numpy(host.rotation_matrices_t(rot, tilt).astype(np.float32)).to(dev).contiguous()
    return kernels.project_slices(dft, rts, whiten=whiten, normalize=whiten, pad=pad).cpu().numpy()


def _write_star(path: str, lines, loop, mrcs_name: str) -> None:
    """Input header, then each particle row with _rlnImageName replaced by
    '<row>@<mrcs_name>', tab-separated (project3d_and_whiten_cuda.py:169-190)."""
    c_img = loop.labels["_rlnImageName"]
    with open(path, "w") as fh:
        fh.writelines(lines[:loop.data_start])
        for r in loop.data_line_indices:
            tok = loop.tokens(r)
            tok[c_img] = f"{r - loop.data_start + 1}@{mrcs_name}"
            fh.write("\t".join(tok) + "\n")


def run_project3d(args: argparse.Namespace, *, volume: Optional[np.ndarray] = None) -> Dict[str, float]:
    """The main of project3d_and_whiten_cuda.py (:39-190): projections of
    ``args.i`` (or ``volume``) at the STAR file's (rot, tilt), psi ignored,
    written to ``<o>.mrcs`` with ``<o>.star`` = the input header and rows,
    each _rlnImageName replaced by ``<row>@<o>.mrcs`` (tab-separated), both
    in the working directory (``<o>`` = the base name of --o without its
    extension).  With --return_array the stack is not written.  Rows after
    the particle loop are not copied."""
    lines, loop, rot, tilt = _read_angles(args.ang)
    vol = _load_volume(volume if volume is not None else args.i)
    dev = _device(getattr(args, "gpuid", None))
    out_base = extract_filename(args.o)
    out_mrcs, out_star = out_base + ".mrcs", out_base + ".star"
    t0 = time.time()
    stack = _project(vol, rot, tilt, not args.skip_padding, bool(args.normal_background_powerspectrum), dev)
    t1 = time.time()
    if not getattr(args, "return_array", False):
        mrc.write_mrc(out_mrcs, stack, voxel_size=0.0)
    _write_star(out_star, lines, loop, out_mrcs)
    t2 = time.time()
    print("Project3D, total execution time = ", round(t2 - t0, 4), " seconds")
    return {"projections": float(len(rot)), "box": float(vol.shape[0]), "gpu_s": t1 - t0, "total_s": t2 - t0}


def run_projection(particle_star, mrc_volume, whitening, do_skip_padding, do_return_array, gpuid,
                   output_filename) -> Optional[np.ndarray]:
    """Library form of the projector (project3d_and_whiten_cuda_nowrite.py:82-238):
    the STAR file goes to ``output_filename`` as given and the stack to
    ``output_filename + '.mrcs'``, which the STAR rows name; with
    ``do_return_array`` the stack is returned as float32 [M, N, N] instead
    of written (the STAR file is written either way).  ``mrc_volume`` is an
    MRC path or a cubic array; ``gpuid`` None means the current HIP device."""
    lines, loop, rot, tilt = _read_angles(particle_star)
    vol = _load_volume(mrc_volume)
    dev = _device(gpuid)
    out_mrcs = str(output_filename) + ".mrcs"
    t0 = time.time()
    stack = _project(vol, rot, tilt, not do_skip_padding, bool(whitening), dev)
    if not do_return_array:
        mrc.write_mrc(out_mrcs, stack, voxel_size=0.0)
    _write_star(str(output_filename), lines, loop, out_mrcs)
    print("Project3D, total execution time = ", round(time.time() - t0, 4), " seconds")
    return stack if do_return_array else None


def main(argv: Optional[Sequence[str]] = None) -> None:
    run_project3d(create_project3d_parser().parse_args(argv))


if __name__ == "__main__":
    main()
