"""Oracle: integrated projector, whitening and the packed template bank.

Test infrastructure only (see ``oracle/__init__.py``).  Follows v606:837-1155
(and its twin project3d_and_whiten_cuda.py:100-291).
"""
from __future__ import annotations

import math

import numpy as np

from .fourier import centered_full_to_packed, pre_shift

PI_REF = 3.14159265359  # the literal used by v606:170-172


def euler_angles_to_matrix(alpha, beta, gamma):
    """v606:168-187 (RELION ZYZ, degrees, pi written as 3.14159265359)."""
    alpha = alpha / 180.0 * PI_REF
    beta = beta / 180.0 * PI_REF
    gamma = gamma / 180.0 * PI_REF
    ca, cb, cg = math.cos(alpha), math.cos(beta), math.cos(gamma)
    sa, sb, sg = math.sin(alpha), math.sin(beta), math.sin(gamma)
    cc, cs, sc, ss = cb * ca, cb * sa, sb * ca, sb * sa
    return [[cg * cc - sg * sa, cg * cs + sg * ca, -cg * sb],
            [-sg * cc - cg * sa, -sg * cs + cg * ca, sg * sb],
            [sc, ss, cb]]


def rotation_matrices_t(rot, tilt) -> np.ndarray:
    """v606:214-215: float32 Euler(rot, tilt, 0), transposed."""
    return np.stack([np.asarray(euler_angles_to_matrix(r, t, 0.0), np.float32).T
                     for r, t in zip(rot, tilt)]).astype(np.float32)


def generate_soft_mask(shape, radius, edge_width) -> np.ndarray:
    """v606:446-457."""
    center = np.asarray(shape) // 2
    y, x = np.indices(shape)
    d = np.sqrt((x - center[1]) ** 2 + (y - center[0]) ** 2)
    mask = np.zeros_like(d)
    mask[d <= radius] = 1.0
    if edge_width > 0:
        edge = (d > radius) & (d <= radius + edge_width)
        mask[edge] = 0.5 * (1.0 + np.cos(np.pi * (d[edge] - radius) / edge_width))
    return mask.astype(np.float32)


def expand_radial_values(values, size) -> np.ndarray:
    """v606:460-468: value at truncated integer radius; 0 beyond the table."""
    v = np.asarray(values, np.float32).reshape(-1)
    c = size // 2
    y, x = np.indices((size, size))
    r = np.sqrt((y - c) ** 2 + (x - c) ** 2).astype(np.int64)
    out = np.zeros((size, size), np.float32)
    ok = r < v.shape[0]
    out[ok] = v[r[ok]]
    return out


def sinc2_weights(n2: int, dtype=np.float32) -> np.ndarray:
    """v606:837-853: sinc(|f|)^2 with f = fftshift(fftfreq(2N)) on each axis."""
    f = np.fft.fftshift(np.fft.fftfreq(n2)).astype(dtype)
    r = np.sqrt(f[:, None, None] ** 2 + f[None, :, None] ** 2 + f[None, None, :] ** 2)
    return (np.sinc(r) ** 2).astype(dtype)


def sinc2_filter(volume, dtype=np.float64) -> np.ndarray:
    """apply_sinc2_filter_inplace_3d (v606:837-853;
    project3d_and_whiten_cuda_nowrite.py:42-81) on any [D, H, W] shape:
    volume * sinc(|f|)^2 with f = fftshift(fftfreq(n)) per axis (out of place)."""
    D, H, W = volume.shape
    fz, fy, fx = (np.fft.fftshift(np.fft.fftfreq(n)) for n in (D, H, W))
    r = np.sqrt(fz[:, None, None] ** 2 + fy[None, :, None] ** 2 + fx[None, None, :] ** 2)
    return (np.asarray(volume, dtype) * np.sinc(r) ** 2).astype(dtype)


def projection_base_grid(image_shape) -> np.ndarray:
    """build_projection_base_grid (v606:856-863): [H, W//2+1, 3] (z, y, x)
    frequencies, z = 0, y = fftshift(fftfreq(H)), x = rfftfreq(W)."""
    _, H, W = image_shape
    yy, xx = np.meshgrid(np.fft.fftshift(np.fft.fftfreq(H)), np.fft.rfftfreq(W), indexing="ij")
    return np.stack((np.zeros_like(yy), yy, xx), -1)


def grid_sample_coordinates(base_zyx: np.ndarray, rts: np.ndarray, image_shape, dft_shape):
    """rotated_grid_to_grid_sample (v606:866-897) in float64: rotate the
    (x, y, z) base vectors by each R, mirror where the rotated x < 0, array
    coordinates (z, y centred at size//2, x from 0), align_corners=True
    normalisation by (dft_shape/2 - 1/2), returned in (x, y, z) order with the
    mirror mask."""
    base_xyz = np.asarray(base_zyx, np.float64)[..., ::-1]
    rot = np.einsum("mij,...j->m...i", np.asarray(rts, np.float64), base_xyz)
    conj = rot[..., 0] < 0
    rot = np.where(conj[..., None], -rot, rot)
    D, H, W = image_shape
    coords = np.stack((rot[..., 2] * D + D // 2, rot[..., 1] * H + H // 2, rot[..., 0] * W), -1)
    norm = coords / (0.5 * np.asarray(dft_shape, np.float64) - 0.5) - 1.0
    return norm[..., ::-1].copy(), conj


def volume_dft(volume, dtype=np.float32, pad=True) -> np.ndarray:
    """v606:1068-1075 / project3d_and_whiten_cuda.py:103-117: pad N/2 (unless
    --skip_padding), sinc^2 over the (padded) shape, fftshift, rfftn,
    fftshift(z, y)."""
    n = volume.shape[0]
    vol = np.pad(np.asarray(volume, dtype), n // 2) if pad else np.asarray(volume, dtype)
    vol = vol * sinc2_weights(vol.shape[0], dtype)
    dft = np.fft.rfftn(np.fft.fftshift(vol))
    return np.fft.fftshift(dft, axes=(0, 1))


def slice_coordinates(rt: np.ndarray, n2: int):
    """v606:856-896: array coordinates (z, y, x) of one central slice and its conj mask."""
    fy = np.fft.fftshift(np.fft.fftfreq(n2)).astype(np.float32)
    fx = np.fft.rfftfreq(n2).astype(np.float32)
    yy, xx = np.meshgrid(fy, fx, indexing="ij")
    base = np.stack((xx, yy, np.zeros_like(xx)), -1).astype(np.float64)  # x, y, z
    rotd = base @ rt.astype(np.float64).T  # R @ v
    gx, gy, gz = rotd[..., 0], rotd[..., 1], rotd[..., 2]
    conj = gx < 0
    sgn = np.where(conj, -1.0, 1.0)
    gx, gy, gz = gx * sgn, gy * sgn, gz * sgn
    return gz * n2 + n2 // 2, gy * n2 + n2 // 2, gx * n2, conj


def trilinear_border(vol: np.ndarray, cz, cy, cx):
    """F.grid_sample semantics: trilinear, padding_mode='border', align_corners=True."""
    D, H, W = vol.shape
    cz = np.clip(cz, 0, D - 1)
    cy = np.clip(cy, 0, H - 1)
    cx = np.clip(cx, 0, W - 1)
    z0 = np.floor(cz).astype(np.int64)
    y0 = np.floor(cy).astype(np.int64)
    x0 = np.floor(cx).astype(np.int64)
    tz, ty, tx = cz - z0, cy - y0, cx - x0
    out = np.zeros(cz.shape, dtype=np.result_type(vol.dtype, np.complex128))
    for dz in (0, 1):
        wz = tz if dz else 1 - tz
        zi = z0 + dz
        for dy in (0, 1):
            wy = ty if dy else 1 - ty
            yi = y0 + dy
            for dx in (0, 1):
                wx = tx if dx else 1 - tx
                xi = x0 + dx
                ok = (zi < D) & (yi < H) & (xi < W)
                v = vol[np.minimum(zi, D - 1), np.minimum(yi, H - 1), np.minimum(xi, W - 1)]
                out += np.where(ok, wz * wy * wx, 0.0) * v
    return out


def central_slice(dft: np.ndarray, rt: np.ndarray, dtype=np.float32) -> np.ndarray:
    """One sampled central slice [2N, N+1] with the Friedel conjugation (v606:1099-1113)."""
    n2 = dft.shape[0]
    cz, cy, cx, conj = slice_coordinates(rt, n2)
    s = trilinear_border(dft, cz, cy, cx)
    s = np.where(conj, np.conj(s), s)
    return s.astype(np.complex64 if dtype == np.float32 else np.complex128)


def slice_to_projection(s: np.ndarray, dtype=np.float32, box=None) -> np.ndarray:
    """v606:1122-1130 / project3d_and_whiten_cuda.py:127-132: ifftshift rows,
    irfftn, ifftshift, crop the padding (box = the unpadded size; default
    half the slice length)."""
    n2 = s.shape[0]
    n = n2 // 2 if box is None else box
    p = np.fft.irfft2(np.fft.ifftshift(s, axes=0), s=(n2, n2), axes=(0, 1))
    p = np.fft.ifftshift(p, axes=(0, 1))
    off = (n2 - n) // 2
    return p[off:off + n, off:off + n].astype(dtype)


def ring_map(n: int) -> np.ndarray:
    """v606:902-905: round(sqrt((y-n//2)^2 + (x-n//2)^2)) (half-to-even)."""
    y, x = np.indices((n, n)).astype(np.float64)
    return np.round(np.sqrt((y - n // 2) ** 2 + (x - n // 2) ** 2)).astype(np.int64)


def divide_by_spectrum(img: np.ndarray, dtype=np.float32) -> np.ndarray:
    """getSpectrum_divideBySpectrum_torch (project3d_and_whiten_cuda.py:196-230;
    the first half of v606:938-992): ring mean of |F| -> divide (DC ring
    untouched, empty rings 1) -> inverse, real part."""
    n = img.shape[0]
    cdt = np.complex64 if dtype == np.float32 else np.complex128
    F = np.fft.fftshift(np.fft.fft2(img.astype(dtype))).astype(cdt)
    rm = ring_map(n)
    nr = int(rm.max()) + 1
    spec = np.zeros(nr, dtype)
    np.add.at(spec, rm.ravel(), np.abs(F).ravel().astype(dtype))
    cnt = np.bincount(rm.ravel(), minlength=nr).astype(dtype)
    div = np.ones(nr, dtype)
    nz = cnt > 0
    div[nz] = cnt[nz] / spec[nz]
    div[0] = 1.0
    F = F * div[rm]
    return np.fft.ifft2(np.fft.ifftshift(F)).real.astype(dtype)


def normalize_plane(img: np.ndarray, dtype=np.float32, eps=1e-12) -> np.ndarray:
    """normalize_torch (project3d_and_whiten_cuda.py:231-291; the second half
    of v606:938-992): least-squares plane with x = row index and y = column
    index (the call-site swap of v606:916-918 / p3d:84-85; a plane fitted
    over the whole grid is the same either way), subtract, then
    (x - mean) / std (population) if std > eps, else x - mean."""
    n = img.shape[0]
    im = img.astype(dtype)
    yy, xx = np.indices((n, n)).astype(dtype)
    cx, cy = yy, xx
    x, y, z = cx.ravel(), cy.ravel(), im.ravel()
    A = np.array([[np.sum(x * x), np.sum(x * y), np.sum(x)],
                  [np.sum(x * y), np.sum(y * y), np.sum(y)],
                  [np.sum(x), np.sum(y), float(x.size)]], np.float64)
    b = np.array([np.sum(x * z), np.sum(y * z), np.sum(z)], np.float64)
    coef = np.linalg.solve(A, b).astype(dtype)
    im = im - (coef[0] * cx + coef[1] * cy + coef[2])
    mean = im.mean(dtype=np.float64)
    std = im.std(dtype=np.float64)
    if std > eps:
        im = (im - mean) / std
    else:
        im = im - mean
    return im.astype(dtype)


def whiten(img: np.ndarray, dtype=np.float32) -> np.ndarray:
    """v606:938-992 (batch == 1 path) / p3d:196-291: divide_by_spectrum then
    normalize_plane."""
    return normalize_plane(divide_by_spectrum(img, dtype), dtype)


def template_from_projection(proj, new_box, mask, packed_fsc, dtype=np.float32):
    """v606:1131-1147: whiten, mask, fft2, centre, crop, pack, FSC, unit L2, pre-shift."""
    n = proj.shape[0]
    im = whiten(proj, dtype)
    if mask is not None:
        im = im * mask.astype(dtype)
    F = np.fft.fftshift(np.fft.fft2(im))
    if new_box != n:
        lo, hi = n // 2 - new_box // 2, n // 2 + new_box // 2
        F = F[lo:hi, lo:hi]
    packed = centered_full_to_packed(F) * packed_fsc.astype(dtype)
    packed = packed / np.linalg.norm(packed.ravel())
    cdt = np.complex64 if dtype == np.float32 else np.complex128
    return pre_shift(packed).astype(cdt)


def build_templates(volume, rot, tilt, new_box, mask, packed_fsc, dtype=np.float32,
                    return_projections=False):
    """IntegratedProjector.build_templates (v606:1053-1155) on arrays.

    Returns the pre-shifted packed bank [M, n, n/2+1]."""
    dft = volume_dft(volume, dtype)
    mats = rotation_matrices_t(rot, tilt)
    out, projs = [], []
    for rt in mats:
        proj = slice_to_projection(central_slice(dft, rt, dtype), dtype)
        projs.append(proj)
        out.append(template_from_projection(proj, new_box, mask, packed_fsc, dtype))
    bank = np.stack(out)
    return (bank, np.stack(projs)) if return_projections else bank


def project3d(volume, rot, tilt, whitening=False, pad=True, dtype=np.float32) -> np.ndarray:
    """The projection loop of project3d_and_whiten_cuda.py main (:103-147):
    [M, N, N] real projections at Euler(rot, tilt, 0), optionally ring-divided
    and plane-normalised (--normal_background_powerspectrum); pad=False is
    --skip_padding."""
    n = volume.shape[0]
    dft = volume_dft(volume, dtype, pad=pad)
    out = []
    for rt in rotation_matrices_t(rot, tilt):
        p = slice_to_projection(central_slice(dft, rt, dtype), dtype, box=n)
        if whitening:
            p = normalize_plane(divide_by_spectrum(p, dtype), dtype)
        out.append(p)
    return np.stack(out)
