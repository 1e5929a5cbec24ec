"""Oracle: the v606 FP32 path in torch on the CPU, with the reference's batching.

Test infrastructure only (see ``oracle/__init__.py``): the CPU baseline that
bench.py times beside the GPU, and nothing on the product path.  The same
operation sequence as the reference, in torch float32 on the host, with its
chunking:
  * templates: pad N/2, sinc^2, rfftn of the padded volume, central slices
    sampled by grid_sample (trilinear, border, align_corners) 8 projections
    per batch, then one orientation at a time: irfft, crop, whitening (ring
    division + plane normalisation), mask, fft2, crop to newbox, pack, FSC,
    unit norm, pre-shift (v606:1053-1155);
  * particles: CTF of a batch of up to 256 records (v606:1735-1745), then per
    particle its psi images rotated by grid_sample (bilinear, zeros), masked,
    fft2, cropped, packed, times the CTF, conjugated, unit norm, pre-shifted
    (v606:1340-1374);
  * search: for each particle, its local templates in chunks of 16 against
    its psi spectra in chunks of 32: product, irfft2, C x C crop, max and
    argmax (v606:1418-1449).
The numpy oracle (oracle.project / oracle.particles / oracle.fourier) pins
each step; tests/test_oracle_torch_cpu.py checks this file against it.
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

from . import fourier as ofr
from . import particles as opt
from . import project as opj


def _pack(centered: torch.Tensor) -> torch.Tensor:
    n = centered.shape[-1]
    rows = torch.from_numpy(ofr.packed_row_index(n))
    cols = torch.from_numpy(ofr.packed_col_index(n))
    return centered.index_select(-2, rows).index_select(-1, cols)


def _pre_shift(packed: torch.Tensor) -> torch.Tensor:
    return torch.fft.fftshift(packed, dim=(-2, -1))


def volume_dft(volume) -> torch.Tensor:
    """Padded, sinc^2-weighted, centred half spectrum [2N, 2N, N+1] (v606:1063-1075)."""
    vol = torch.as_tensor(np.asarray(volume, np.float32))
    n = vol.shape[0]
    vol = F.pad(vol, [n // 2] * 6)
    vol = vol * torch.from_numpy(opj.sinc2_weights(vol.shape[0], np.float32))
    dft = torch.fft.rfftn(torch.fft.fftshift(vol), dim=(-3, -2, -1))
    return torch.fft.fftshift(dft, dim=(-3, -2))


def _slice_grid(rt: np.ndarray, n2: int):
    """grid_sample coordinates (x, y, z in [-1, 1]) of one central slice and its conj mask."""
    cz, cy, cx, conj = opj.slice_coordinates(rt, n2)
    w = n2 // 2 + 1
    g = np.stack((2.0 * cx / (w - 1) - 1.0, 2.0 * cy / (n2 - 1) - 1.0, 2.0 * cz / (n2 - 1) - 1.0), -1)
    return torch.from_numpy(g.astype(np.float32)), torch.from_numpy(conj)


class Whitening:
    """Ring division by the amplitude spectrum, then least-squares plane
    removal and standardisation (v606:902-992 / p3d:196-291)."""

    def __init__(self, n: int):
        self.n = n
        rm = opj.ring_map(n)
        self.rings = torch.from_numpy(rm.reshape(-1))
        self.nr = int(rm.max()) + 1
        self.count = torch.bincount(self.rings, minlength=self.nr).to(torch.float32)
        yy, xx = np.indices((n, n)).astype(np.float32)
        self.cx, self.cy = torch.from_numpy(yy), torch.from_numpy(xx)
        x, y = yy.ravel().astype(np.float64), xx.ravel().astype(np.float64)
        self.A = torch.tensor([[np.sum(x * x), np.sum(x * y), np.sum(x)], [np.sum(x * y), np.sum(y * y), np.sum(y)],
                               [np.sum(x), np.sum(y), float(x.size)]], dtype=torch.float64)

    def __call__(self, img: torch.Tensor) -> torch.Tensor:
        n = self.n
        Fc = torch.fft.fftshift(torch.fft.fft2(img))
        spec = torch.zeros(self.nr, dtype=torch.float32).index_add_(0, self.rings, Fc.abs().reshape(-1))
        div = torch.ones(self.nr, dtype=torch.float32)
        nz = self.count > 0
        div[nz] = self.count[nz] / spec[nz]
        div[0] = 1.0
        Fc = Fc * div[self.rings].reshape(n, n)
        im = torch.fft.ifft2(torch.fft.ifftshift(Fc)).real.to(torch.float32)
        z = im.reshape(-1).to(torch.float64)
        x, y = self.cx.reshape(-1).to(torch.float64), self.cy.reshape(-1).to(torch.float64)
        b = torch.stack([(x * z).sum(), (y * z).sum(), z.sum()])
        coef = torch.linalg.solve(self.A, b).to(torch.float32)
        im = im - (coef[0] * self.cx + coef[1] * self.cy + coef[2])
        mean = im.to(torch.float64).mean()
        std = im.to(torch.float64).std(unbiased=False)
        im = (im - mean) / std if std > 1e-12 else im - mean
        return im.to(torch.float32)


def build_templates(volume, rot, tilt, new_box: int, mask, packed_fsc, batch: int = 8, dft=None) -> torch.Tensor:
    """Pre-shifted packed bank [M, n, n/2+1] complex64 (v606:1053-1155);
    ``dft`` = volume_dft(volume) when already computed."""
    N = np.asarray(volume).shape[0]
    dft = volume_dft(volume) if dft is None else dft
    n2 = dft.shape[0]
    dft_real = torch.view_as_real(dft).permute(3, 0, 1, 2).unsqueeze(0).contiguous()
    mats = opj.rotation_matrices_t(rot, tilt)
    white = Whitening(N)
    maskt = None if mask is None else torch.as_tensor(np.asarray(mask, np.float32))
    fsc = torch.as_tensor(np.asarray(packed_fsc, np.float32))
    lo, hi = N // 2 - new_box // 2, N // 2 + new_box // 2
    pad = N // 2
    out = torch.empty((len(mats), new_box, new_box // 2 + 1), dtype=torch.complex64)
    for s in range(0, len(mats), batch):
        grids, conjs = zip(*(_slice_grid(rt, n2) for rt in mats[s:s + batch]))
        grid = torch.stack(grids).unsqueeze(0)               # [1, k, 2N, N+1, 3]
        sampled = F.grid_sample(dft_real, grid, mode="bilinear", padding_mode="border", align_corners=True)
        sc = torch.view_as_complex(sampled.squeeze(0).permute(1, 2, 3, 0).contiguous())
        sc = torch.where(torch.stack(conjs), sc.conj(), sc)
        for k in range(sc.shape[0]):                          # per-orientation tail
            proj = torch.fft.irfftn(torch.fft.ifftshift(sc[k:k + 1], dim=(-2,)), dim=(-2, -1))
            proj = torch.fft.ifftshift(proj, dim=(-2, -1))[0, pad:pad + N, pad:pad + N].to(torch.float32)
            im = white(proj)
            if maskt is not None:
                im = im * maskt
            Fc = torch.fft.fftshift(torch.fft.fft2(im))
            if new_box != N:
                Fc = Fc[lo:hi, lo:hi]
            packed = _pack(Fc) * fsc
            packed = packed / torch.linalg.vector_norm(packed.reshape(-1))
            out[s + k] = _pre_shift(packed)
    return out


def ctf_batch(ctf: opt.PackedCTF, defocus) -> torch.Tensor:
    """Packed CTFs of a batch of records [k, n, n/2+1] float32."""
    return torch.from_numpy(np.stack([ctf.one(*d) for d in defocus]).astype(np.float32))


def _rotation_grid(N: int, angle: float) -> torch.Tensor:
    sy, sx = opt.rotation_source(N, angle)
    g = np.stack((2.0 * sx / (N - 1) - 1.0, 2.0 * sy / (N - 1) - 1.0), -1)
    return torch.from_numpy(g.astype(np.float32))


def prepare_particle(img, eff_psi, new_box: int, mask, ctf: torch.Tensor) -> torch.Tensor:
    """Pre-shifted packed psi spectra of one particle [P, n, n/2+1] (v606:1340-1374)."""
    x = torch.as_tensor(np.asarray(img, np.float32))
    N = x.shape[0]
    grid = torch.stack([_rotation_grid(N, a) for a in eff_psi])
    r = F.grid_sample(x.expand(len(eff_psi), 1, N, N), grid, mode="bilinear", padding_mode="zeros",
                      align_corners=True)[:, 0]
    if mask is not None:
        r = r * torch.as_tensor(np.asarray(mask, np.float32))
    Fc = torch.fft.fftshift(torch.fft.fft2(r), dim=(-2, -1))
    if new_box != N:
        lo, hi = N // 2 - new_box // 2, N // 2 + new_box // 2
        Fc = Fc[:, lo:hi, lo:hi]
    p = torch.conj(_pack(Fc) * ctf)
    p = p / torch.linalg.vector_norm(p.reshape(p.shape[0], -1), dim=1)[:, None, None]
    return _pre_shift(p).to(torch.complex64)


def correlate(templates: torch.Tensor, spectra: torch.Tensor, crop: int, chunk_a: int = 16, chunk_b: int = 32):
    """(scores, y, x) [A, B] of correlate_templates_and_particle (v606:1418-1449)."""
    n = templates.shape[-2]
    idx = torch.from_numpy(ofr.crop_indices(n, crop))
    A, B = templates.shape[0], spectra.shape[0]
    scores = torch.empty((A, B), dtype=torch.float32)
    ys = torch.empty((A, B), dtype=torch.int64)
    xs = torch.empty((A, B), dtype=torch.int64)
    mult = float(1024 * n * n)
    for i in range(0, A, chunk_a):
        ta = templates[i:i + chunk_a]
        for j in range(0, B, chunk_b):
            pb = spectra[j:j + chunk_b]
            ccg = torch.fft.irfft2(ta[:, None] * pb[None], s=(n, n), dim=(-2, -1))
            cr = ccg.index_select(-2, idx).index_select(-1, idx)
            v, k = cr.reshape(cr.shape[0], cr.shape[1], -1).max(dim=-1)
            scores[i:i + chunk_a, j:j + chunk_b] = v * mult
            ys[i:i + chunk_a, j:j + chunk_b] = k // crop
            xs[i:i + chunk_a, j:j + chunk_b] = k % crop
    return scores, ys, xs


def search(bank: torch.Tensor, images, defocus, eff_psi, local, new_box: int, mask, ctf: opt.PackedCTF,
           crop: int, ctf_batch_size: int = 256):
    """The per-particle loop of run_search (v606:1735-1790): per batch of up to
    256 particles their CTFs, then per particle its psi spectra and the
    chunked correlation over its local templates.  Returns per particle
    (peak score, flat best pair, y, x) and the number of scores."""
    out, scores = [], 0
    for s in range(0, len(images), ctf_batch_size):
        ctfs = ctf_batch(ctf, defocus[s:s + ctf_batch_size])
        for k in range(ctfs.shape[0]):
            b = s + k
            idx = local[b]
            if len(idx) == 0:
                out.append(None)
                continue
            spec = prepare_particle(images[b], eff_psi[b], new_box, mask, ctfs[k])
            sc, ys, xs = correlate(bank[torch.as_tensor(np.asarray(idx, np.int64))], spec, crop)
            flat = int(torch.argmax(sc.reshape(-1)))
            out.append((float(sc.reshape(-1)[flat]), flat, int(ys.reshape(-1)[flat]), int(xs.reshape(-1)[flat])))
            scores += sc.numel()
    return out, scores
