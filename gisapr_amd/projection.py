"""The grid_sample projector's building blocks, under the names the batched
search engine resolves from the search module's namespace (v606:837-1031):

* ``apply_sinc2_filter_inplace_3d(volume, z_chunk=8)`` -- sinc^2(|f|) window
  in place (v606:837-853), one HIP pass (``gis_sinc2_filter``)
* ``build_projection_base_grid(image_shape, device)`` -- the (0, fy, fx)
  frequency plane of an rfft slice (v606:856-863), a host-built table
* ``rotated_grid_to_grid_sample(base, R, image_shape, dft_shape)`` -- rotated
  F.grid_sample coordinates + Friedel mask (v606:866-897), one HIP pass
  (``gis_slice_sample_grid``)
* ``ProjectionWhitening(size, device)`` -- ring-mean spectrum division, plane
  removal and standardisation of [B, H, W] projections (v606:899-1031), the
  template kernel's image mode (``gis_whiten_images``)

The search path itself does not use them (the template kernel samples,
whitens and packs in one launch); they serve callers that assemble their own
grid_sample projector.  Device tensors only: there is no CPU path.
"""
from __future__ import annotations

from typing import Tuple

import numpy as np
import torch

from . import kernels

__all__ = ["apply_sinc2_filter_inplace_3d", "build_projection_base_grid", "rotated_grid_to_grid_sample",
           "ProjectionWhitening"]


def apply_sinc2_filter_inplace_3d(volume: torch.Tensor, z_chunk: int = 8) -> None:
    """volume[z, y, x] *= sinc(|f|)^2 in place, f = fftshift(fftfreq(n)) per
    axis (v606:837-853).  ``z_chunk`` is accepted for the signature: the kernel
    computes the window per voxel and never materialises a frequency grid.
    The volume must be a contiguous float32 tensor on the HIP device."""
    if not isinstance(volume, torch.Tensor) or volume.ndim != 3:
        raise ValueError("volume must be 3-D")
    if int(z_chunk) <= 0:
        raise ValueError("z_chunk must be positive")
    kernels.sinc2_filter_(volume)


def _shifted_freq(n: int) -> np.ndarray:
    # torch: fftshift(arange(n) * (1/n) in float32) -> (i - n//2) * float32(1/n)
    return (np.arange(n, dtype=np.float32) - np.float32(n // 2)) * np.float32(1.0 / n)


def build_projection_base_grid(image_shape: Tuple[int, int, int], device) -> torch.Tensor:
    """[H, W//2+1, 3] float32 (z, y, x) frequencies of the unrotated rfft
    slice: z = 0, y = fftshift(fftfreq(H)), x = rfftfreq(W) (v606:856-863)."""
    _, height, width = (int(v) for v in image_shape)
    fy = _shifted_freq(height)
    fx = np.arange(width // 2 + 1, dtype=np.float32) * np.float32(1.0 / width)
    grid = np.zeros((height, width // 2 + 1, 3), np.float32)
    grid[..., 1] = fy[:, None]
    grid[..., 2] = fx[None, :]
    return torch.from_numpy(grid).to(device)


def rotated_grid_to_grid_sample(base_grid_zyx: torch.Tensor, rotation_matrices_xyz: torch.Tensor,
                                image_shape: Tuple[int, int, int],
                                dft_shape: Tuple[int, int, int]) -> Tuple[torch.Tensor, torch.Tensor]:
    """(grid [M, H, W, 3] in grid_sample's x, y, z order, conjugate mask
    [M, H, W]) for rotations acting on xyz vectors (v606:866-897): rotate,
    mirror the points whose rotated x frequency is negative, centre z and y,
    normalise by (dft_shape / 2 - 1/2) for align_corners=True."""
    base = base_grid_zyx
    R = torch.as_tensor(rotation_matrices_xyz)
    if R.device != base.device or R.dtype != torch.float32:
        R = R.to(base.device, torch.float32)
    return kernels.slice_sample_grid(base, R.reshape(-1, 3, 3), image_shape, dft_shape)


class ProjectionWhitening:
    """Whitening of real projections (v606:899-1031): divide the centred
    spectrum by its ring mean of |F| (DC ring kept), inverse transform, remove
    the least-squares plane, then (x - mean) / std (x - mean when std <= 1e-12).

    The ring map, counts and plane-fit grids are kept as attributes, as the
    reference keeps them; the arithmetic runs in ``gis_whiten_images``."""

    def __init__(self, size: int, device) -> None:
        self.size = int(size)
        if self.size not in kernels.SUPPORTED_BOXES:
            raise ValueError(f"projection size {self.size} not in {kernels.SUPPORTED_BOXES}")
        self.device = torch.device(device)
        n, c = self.size, self.size // 2
        idx = np.arange(n, dtype=np.float32)
        yy, xx = np.meshgrid(idx, idx, indexing="ij")
        ring = np.round(np.sqrt((yy - c) ** 2 + (xx - c) ** 2)).astype(np.int64)
        self.radius_map = torch.from_numpy(ring).to(self.device)
        self.radius_flat = self.radius_map.reshape(-1)
        self.max_radius = int(ring.max()) + 1
        self.radius_counts = torch.from_numpy(np.bincount(ring.ravel(), minlength=self.max_radius)
                                              .astype(np.float32)).to(self.device)
        # the reference's plane fit calls the row grid x and the column grid y
        self.coord_x = torch.from_numpy(yy).to(self.device)
        self.coord_y = torch.from_numpy(xx).to(self.device)
        sx, sy = float(yy.sum(dtype=np.float64)), float(xx.sum(dtype=np.float64))
        sxx, syy = float((yy.astype(np.float64) ** 2).sum()), float((xx.astype(np.float64) ** 2).sum())
        sxy = float((yy.astype(np.float64) * xx).sum())
        self.plane_matrix = torch.tensor([[sxx, sxy, sx], [sxy, syy, sy], [sx, sy, float(n * n)]],
                                         dtype=torch.float32, device=self.device)

    def __call__(self, images: torch.Tensor) -> torch.Tensor:
        if images.ndim != 3:
            raise ValueError("Projection whitening expects (B,H,W)")
        if tuple(images.shape[1:]) != (self.size, self.size):
            raise ValueError(f"projections must be {self.size} x {self.size}")
        return kernels.whiten_images(images.to(torch.float32).contiguous(), whiten=True, normalize=True, eps=1e-12)
