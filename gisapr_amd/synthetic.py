"""Seeded synthetic workloads for the candidate-scoring path.

Builds, from fixed seeds, everything one search needs: a random-walk atomic
model split into chains, a HEALPix direction table in the RELION STAR
convention, particle images projected from the model at random orientations
(then CTF, integer shift, white noise at a given SNR and whitening), the
particle metadata (priors and defocus) and a candidate rigid-body transform.

Seeds: atoms 0, orientations 1, CTF 2, shifts 3, noise 4, candidates 5,
arrivals 6.

Particle images are made from the atoms directly (each atom's radial
exponential integrated along the beam, splatted in 2-D), not through the
search's own projector, so that the search has an independent target.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from .density import ELEMENT_Z, splat_parameters

ELEMENT_MIX = (("C", 0.63), ("N", 0.17), ("O", 0.19), ("S", 0.01))

SEED_ATOMS, SEED_ORIENT, SEED_CTF, SEED_SHIFT, SEED_NOISE, SEED_CAND, SEED_ARRIVAL = range(7)


# ---------------------------------------------------------------------------
# Atomic model
# ---------------------------------------------------------------------------

def random_walk_chains(chain_sizes, seed=SEED_ATOMS, step=3.8, confine=40.0):
    """Random walks with ``step`` A steps inside a ball of radius ``confine``.

    Directions are isotropic; a step that would leave the ball is redrawn.
    Returns (coords [A,3] float32 centred at the origin, elements [A], chains [A]).
    """
    rng = np.random.default_rng(seed)
    coords, chains = [], []
    for ci, n in enumerate(chain_sizes):
        p = rng.uniform(-0.5, 0.5, 3) * confine
        for _ in range(int(n)):
            for _try in range(64):
                d = rng.standard_normal(3)
                q = p + step * d / np.linalg.norm(d)
                if np.linalg.norm(q) <= confine:
                    break
            p = q
            coords.append(p.copy())
            chains.append(chr(ord("A") + ci))
    coords = np.asarray(coords, np.float64)
    coords -= coords.mean(0)
    names = [e for e, _ in ELEMENT_MIX]
    probs = np.asarray([w for _, w in ELEMENT_MIX])
    elements = rng.choice(names, size=len(coords), p=probs / probs.sum())
    return coords.astype(np.float32), np.asarray(elements), np.asarray(chains)


def pdb_text(coords, elements, chains) -> str:
    """ATOM records (columns per the PDB format) for the model."""
    lines = []
    for i, ((x, y, z), e, c) in enumerate(zip(coords, elements, chains)):
        name = f" {e:<3s}"
        lines.append(f"ATOM  {(i + 1) % 100000:5d} {name:4s} ALA {c}{(i // 4 + 1) % 10000:4d}    "
                     f"{x:8.3f}{y:8.3f}{z:8.3f}  1.00  0.00          {e:>2s}\n")
    lines.append("END\n")
    return "".join(lines)


# ---------------------------------------------------------------------------
# Directions and orientations
# ---------------------------------------------------------------------------

_JRLL = (2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4)
_JPLL = (1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7)


def _deinterleave(v: int, bits: int):
    x = y = 0
    for b in range(bits):
        x |= ((v >> (2 * b)) & 1) << b
        y |= ((v >> (2 * b + 1)) & 1) << b
    return x, y


def healpix_nest_directions(order: int):
    """HEALPix NESTED pixel centres (Gorski et al. 2005, pix2ang_nest).

    Returns (rot, tilt) in degrees as the reference's STAR writer stores them:
    rot = fmod(phi + 360, 360) - 180, tilt = fmod(theta + 360, 360).
    """
    nside = 1 << order
    npface = nside * nside
    npix = 12 * npface
    fact2 = 4.0 / npix
    fact1 = (nside << 1) * fact2
    nl4 = 4 * nside
    rot, tilt = [], []
    for pix in range(npix):
        face = pix // npface
        ix, iy = _deinterleave(pix % npface, order)
        jr = (_JRLL[face] << order) - ix - iy - 1
        if jr < nside:
            nr = jr
            z = 1 - nr * nr * fact2
            kshift = 0
        elif jr > 3 * nside:
            nr = nl4 - jr
            z = nr * nr * fact2 - 1
            kshift = 0
        else:
            nr = nside
            z = (2 * nside - jr) * fact1
            kshift = (jr - nside) & 1
        jp = (_JPLL[face] * nr + ix - iy + 1 + kshift) // 2
        if jp > nl4:
            jp -= nl4
        if jp < 1:
            jp += nl4
        phi = (jp - (kshift + 1) * 0.5) * (math.pi / 2 / nr)
        r = math.degrees(phi)
        t = math.degrees(math.acos(max(-1.0, min(1.0, z))))
        rot.append(math.fmod(r + 360.0, 360.0) - 180.0)
        tilt.append(math.fmod(t + 360.0, 360.0))
    return np.asarray(rot), np.asarray(tilt)


def random_orientations(n, seed=SEED_ORIENT):
    """Uniform on SO(3): rot U[-180,180), tilt = arccos U[-1,1], psi U[0,360)."""
    rng = np.random.default_rng(seed)
    rot = rng.uniform(-180.0, 180.0, n)
    tilt = np.degrees(np.arccos(rng.uniform(-1.0, 1.0, n)))
    psi = rng.uniform(0.0, 360.0, n)
    return rot, tilt, psi


def euler_zyz(rot, tilt, psi=0.0):
    """RELION ZYZ Euler matrix (degrees), the convention of the search's projector."""
    a, b, g = (math.radians(v) for v in (rot, tilt, psi))
    ca, cb, cg, sa, sb, sg = math.cos(a), math.cos(b), math.cos(g), math.sin(a), math.sin(b), math.sin(g)
    cc, cs, sc, ss = cb * ca, cb * sa, sb * ca, sb * sa
    return np.asarray([[cg * cc - sg * sa, cg * cs + sg * ca, -cg * sb],
                       [-sg * cc - cg * sa, -sg * cs + cg * ca, sg * sb],
                       [sc, ss, cb]])


# ---------------------------------------------------------------------------
# Particle images
# ---------------------------------------------------------------------------

def projected_profile(width, rp, step=0.01):
    """Line integral of exp(-rp * r) along the beam over |w| <= width (voxels)."""
    rho = np.arange(0.0, width * 1.5 + 1.0, step)
    w = np.linspace(-width, width, 2001)
    r = np.sqrt(rho[:, None] ** 2 + w[None, :] ** 2)
    vals = np.trapezoid(np.exp(-rp * r), w, axis=1)
    return rho, vals


def project_atoms(coords, numbers, rot, tilt, psi, shifts, box, apix, res, device="cpu",
                  batch=16):
    """Projection images [B, box, box] of the atom model (float32 torch tensor)."""
    import torch
    width, rp, kn = splat_parameters(apix, res)
    rho, prof = projected_profile(width, rp)
    prof_t = torch.as_tensor(prof, dtype=torch.float32, device=device)
    step = float(rho[1] - rho[0])
    a = torch.as_tensor(np.asarray(coords, np.float64) / apix, dtype=torch.float64, device=device)
    zn = torch.as_tensor(np.asarray(numbers, np.float64) * kn, dtype=torch.float32, device=device)
    keep = zn != 0
    a, zn = a[keep], zn[keep]
    off = torch.arange(-width, width + 1, device=device)
    oy, ox = torch.meshgrid(off, off, indexing="ij")
    oy, ox = oy.reshape(-1), ox.reshape(-1)
    B = len(rot)
    out = torch.zeros((B, box * box), dtype=torch.float32, device=device)
    c = box // 2
    for s in range(0, B, batch):
        e = min(B, s + batch)
        E = torch.as_tensor(np.stack([euler_zyz(rot[i], tilt[i], 0.0) for i in range(s, e)]),
                            dtype=torch.float64, device=device)
        b = torch.einsum("bij,aj->bai", E, a)  # (u, v, w) per atom
        u, v = b[..., 0], b[..., 1]
        th = torch.as_tensor(np.radians(np.asarray(psi[s:e], np.float64)), device=device)
        ct, st = torch.cos(th)[:, None], torch.sin(th)[:, None]
        row = ct * v - st * u + c + torch.as_tensor(np.asarray(shifts[s:e, 0], np.float64), device=device)[:, None]
        col = st * v + ct * u + c + torch.as_tensor(np.asarray(shifts[s:e, 1], np.float64), device=device)[:, None]
        r0 = torch.round(row).long()
        c0 = torch.round(col).long()
        pr = r0[..., None] + oy
        pc = c0[..., None] + ox
        d = torch.sqrt((pr - row[..., None]) ** 2 + (pc - col[..., None]) ** 2)
        fi = torch.clamp((d / step).long(), max=len(prof) - 2)
        frac = (d / step - fi).float()
        val = (prof_t[fi] * (1 - frac) + prof_t[fi + 1] * frac) * zn[None, :, None]
        ok = (pr >= 0) & (pr < box) & (pc >= 0) & (pc < box)
        bidx = torch.arange(e - s, device=device)[:, None, None].expand_as(pr)
        flat = (bidx * box * box + pr * box + pc)[ok]
        out[s:e].view(-1).index_add_(0, flat, val[ok])
    return out.view(B, box, box)


def ctf_physical(box, apix, dfu, dfv, dfa, voltage=300.0, cs=2.7, amp=0.1, device="cpu"):
    """-sin(gamma) on the rfft2 grid [B, box, box//2+1] (no weighting)."""
    import torch
    kv = voltage * 1e3
    lam = 12.2643247 / math.sqrt(kv * (1.0 + kv * 0.978466e-6))
    k1 = math.pi * lam
    k2 = math.pi / 2.0 * cs * 1e7 * lam ** 3
    k3 = math.atan(amp / math.sqrt(1.0 - amp * amp))
    fy = torch.fft.fftfreq(box, d=apix, device=device, dtype=torch.float64)
    fx = torch.fft.rfftfreq(box, d=apix, device=device, dtype=torch.float64)
    y, x = torch.meshgrid(fy, fx, indexing="ij")
    dfu = torch.as_tensor(dfu, dtype=torch.float64, device=device).view(-1, 1, 1)
    dfv = torch.as_tensor(dfv, dtype=torch.float64, device=device).view(-1, 1, 1)
    ang = torch.deg2rad(torch.as_tensor(dfa, dtype=torch.float64, device=device)).view(-1, 1, 1)
    s, c = torch.sin(ang), torch.cos(ang)
    axx = -dfu * c * c - dfv * s * s
    ayy = -dfu * s * s - dfv * c * c
    axy = (dfv - dfu) * s * c
    u2 = x * x + y * y
    gamma = k1 * (axx * x * x + 2 * axy * x * y + ayy * y * y) + k2 * u2 * u2 - k3
    return (-torch.sin(gamma)).float()


def whiten_batch(images):
    """Ring-mean amplitude whitening, least-squares plane removal and
    zero-mean unit-variance scaling, per image (the particle preparation the
    reference applies before searching)."""
    import torch
    B, n, _ = images.shape
    dev = images.device
    yy, xx = torch.meshgrid(torch.arange(n, device=dev, dtype=torch.float64),
                            torch.arange(n, device=dev, dtype=torch.float64), indexing="ij")
    ring = torch.round(torch.sqrt((yy - n // 2) ** 2 + (xx - n // 2) ** 2)).long().view(-1)
    nr = int(ring.max()) + 1
    cnt = torch.bincount(ring, minlength=nr).double()
    F = torch.fft.fftshift(torch.fft.fft2(images.double()), dim=(-2, -1))
    spec = torch.zeros((B, nr), dtype=torch.float64, device=dev)
    spec.scatter_add_(1, ring.view(1, -1).expand(B, -1), F.abs().view(B, -1))
    div = cnt.view(1, -1) / spec.clamp_min(1e-300)
    div[:, 0] = 1.0
    F = F * div[:, ring].view(B, n, n)
    im = torch.fft.ifft2(torch.fft.ifftshift(F, dim=(-2, -1))).real
    A = torch.stack([torch.stack([(xx * xx).sum(), (xx * yy).sum(), xx.sum()]),
                     torch.stack([(xx * yy).sum(), (yy * yy).sum(), yy.sum()]),
                     torch.stack([xx.sum(), yy.sum(), torch.tensor(float(n * n), device=dev, dtype=torch.float64)])])
    b = torch.stack([(im * xx).sum((-2, -1)), (im * yy).sum((-2, -1)), im.sum((-2, -1))], 1)
    coef = torch.linalg.solve(A.cpu(), b.cpu().T).T.to(dev)
    im = im - (coef[:, 0, None, None] * xx + coef[:, 1, None, None] * yy + coef[:, 2, None, None])
    im = im - im.mean((-2, -1), keepdim=True)
    sd = im.std((-2, -1), correction=0, keepdim=True)
    return (im / torch.where(sd > 0, sd, torch.ones_like(sd))).float()


def fsc_logistic(box, apix, crossing_resolution=4.0, width=2.0, threshold=0.143):
    """FSC(k) = 1/(1+exp((k-k0)/w)) with FSC = threshold at shell box*apix/res."""
    k_cross = box * apix / crossing_resolution
    k0 = k_cross - width * math.log(1.0 / threshold - 1.0)
    k = np.arange(box // 2 + 1, dtype=np.float64)
    return 1.0 / (1.0 + np.exp((k - k0) / width))


# ---------------------------------------------------------------------------
# Whole workloads
# ---------------------------------------------------------------------------

@dataclass
class Workload:
    name: str
    box: int
    apix: float
    res: float
    chain_sizes: List[int]
    moving_chain: str
    n_particles: int
    snr: float
    healpix_order: int
    local_range: float
    psi_step: float
    crop: int
    kk: float = 3.0
    mask_radius: int = 110
    mask_edge: int = 6
    fsc_resolution: Optional[float] = None
    confine: float = 40.0
    max_shift: int = 3
    cand_rot_deg: float = 3.0
    cand_trans_A: float = 1.0
    voltage: float = 300.0
    cs: float = 2.7


WORKLOADS: Dict[str, Workload] = {
    # Plain method, CPU-runnable: box 96, 1.5 A/px, 200 particles.
    "box96": Workload("box96", 96, 1.5, 3.0, [1500, 600], "B", 200, 0.1, 2, 15.0, 7.5, 10,
                      mask_radius=44, confine=40.0),
    # Plain FP32 FFT search: box 256, 1.06 A/px, 12,000 atoms, 5,000 particles.
    # Same image generator as box96 (shifts +/-3 px); the search covers +/-8 px (crop 16).
    "box256": Workload("box256", 256, 1.06, 2.12, [4500, 4500, 3000], "C", 5000, 0.05, 3, 10.0,
                       3.75, 16, fsc_resolution=4.0, confine=80.0),
    # Batched PSO step: box 192, 1.3 A/px, 20,000 atoms (moving chain 4,500), 20,000 particles.
    "box192": Workload("box192", 192, 1.3, 2.6, [7750, 7750, 4500], "C", 20000, 0.05, 3, 10.0,
                       3.75, 16, fsc_resolution=4.0, confine=90.0, mask_radius=84),
    # Serving datasets: 12,000-atom complex with a 3,000-atom moving chain at three boxes.
    "serve128": Workload("serve128", 128, 2.0, 4.0, [4500, 4500, 3000], "C", 2000, 0.05, 3, 10.0,
                         3.75, 16, fsc_resolution=8.0, confine=80.0, mask_radius=56),
    "serve192": Workload("serve192", 192, 1.3, 2.6, [4500, 4500, 3000], "C", 10000, 0.05, 3, 10.0,
                         3.75, 16, fsc_resolution=4.0, confine=80.0, mask_radius=84),
    "serve256": Workload("serve256", 256, 1.06, 2.12, [4500, 4500, 3000], "C", 20000, 0.05, 3, 10.0,
                         3.75, 16, fsc_resolution=4.0, confine=80.0),
    # Small fast cases for tests.
    "box64": Workload("box64", 64, 2.0, 4.0, [500, 200], "B", 24, 0.1, 2, 15.0, 7.5, 8,
                      mask_radius=28, confine=28.0),
    "box32": Workload("box32", 32, 3.0, 6.0, [150, 60], "B", 6, 1.0, 1, 30.0, 15.0, 8,
                      mask_radius=13, mask_edge=3, confine=20.0),
}


def candidate_pose(seed=SEED_CAND, rot_deg=3.0, trans_A=1.0):
    """Rotation vector of ``rot_deg`` about a seeded unit axis + ``trans_A`` along a seeded unit vector."""
    rng = np.random.default_rng(seed)
    ax = rng.standard_normal(3)
    ax /= np.linalg.norm(ax)
    tr = rng.standard_normal(3)
    tr /= np.linalg.norm(tr)
    return np.concatenate((ax * rot_deg, tr * trans_A))


def make_workload(name_or_wl, n_particles=None, device="cpu"):
    """Return a dict with the model, directions, particle images and metadata."""
    import torch
    wl = WORKLOADS[name_or_wl] if isinstance(name_or_wl, str) else name_or_wl
    B = wl.n_particles if n_particles is None else int(n_particles)
    coords, elements, chains = random_walk_chains(wl.chain_sizes, confine=wl.confine)
    numbers = np.asarray([ELEMENT_Z.get(e, 0.0) for e in elements])
    drot, dtilt = healpix_nest_directions(wl.healpix_order)
    rot, tilt, psi = random_orientations(B)
    rng = np.random.default_rng(SEED_CTF)
    dfu = rng.uniform(10000.0, 25000.0, B)
    dfv = rng.uniform(10000.0, 25000.0, B)
    dfa = rng.uniform(0.0, 180.0, B)
    shifts = np.random.default_rng(SEED_SHIFT).integers(-wl.max_shift, wl.max_shift + 1, (B, 2))
    img = torch.empty((B, wl.box, wl.box), dtype=torch.float32, device=device)
    gen = torch.Generator(device="cpu").manual_seed(SEED_NOISE)
    chunk = max(1, min(B, (64 << 20) // (wl.box * wl.box * 8)))
    for s in range(0, B, chunk):
        e = min(B, s + chunk)
        sl = slice(s, e)
        im = project_atoms(coords, numbers, rot[sl], tilt[sl], psi[sl], shifts[sl], wl.box, wl.apix,
                           wl.res, device)
        ctf = ctf_physical(wl.box, wl.apix, dfu[sl], dfv[sl], dfa[sl], wl.voltage, wl.cs, 0.1, device)
        im = torch.fft.irfft2(torch.fft.rfft2(im) * ctf, s=(wl.box, wl.box))
        var = im.var(dim=(-2, -1), keepdim=True)
        noise = torch.randn(im.shape, generator=gen, dtype=torch.float32).to(device)
        im = im + noise * torch.sqrt(var / wl.snr)
        img[sl] = whiten_batch(im)
    fsc = None if wl.fsc_resolution is None else fsc_logistic(wl.box, wl.apix, wl.fsc_resolution)
    return {
        "workload": wl, "coords": coords, "elements": elements, "chains": chains,
        "numbers": numbers, "dir_rot": drot, "dir_tilt": dtilt,
        "images": img, "rot": rot, "tilt": tilt, "psi": psi,
        "dfu": dfu, "dfv": dfv, "dfa": dfa, "shifts": shifts, "fsc": fsc,
        "pose": candidate_pose(rot_deg=wl.cand_rot_deg, trans_A=wl.cand_trans_A),
    }
