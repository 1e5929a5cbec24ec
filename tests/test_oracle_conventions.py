"""Conventions of the path pinned by hand values and by physics, not by
restating the formulas: the RELION ZYZ Euler matrix (v606:168-187), the ring
index of the whitening (v606:902-905), and recovery of the synthetic truth
by the search on noiseless particles placed on the direction and psi grids.

On the shift: v606 correlates pre-shifted operands (fftshift over both axes,
v606:535-537), which rolls the n/2+1 packed columns by (n/2+1)//2 and so is
not the textbook cross-correlation (correlate_templates_and_particle,
v606:1418-1449); its reported shift is not the particle's physical shift.
The textbook correlation of the same prepared operands (the roll undone)
recovers the physical shift in the frame the particle is searched in: the
lag is -R(-psi) s for a particle shifted by s = (row, column) after the
in-plane rotation psi."""
import math

import numpy as np
import pytest

from gisapr_amd import host, synthetic
from oracle import particles as opt
from oracle import project as opj
from oracle import search as osr
from oracle import splat as osp

C30, S30 = math.sqrt(3.0) / 2.0, 0.5
HAND = [((90.0, 0.0, 0.0), [[0, 1, 0], [-1, 0, 0], [0, 0, 1]]),
        ((-30.0, 0.0, 0.0), [[C30, -S30, 0], [S30, C30, 0], [0, 0, 1]]),
        ((0.0, 90.0, 0.0), [[0, 0, -1], [0, 1, 0], [1, 0, 0]]),
        ((0.0, 0.0, 90.0), [[0, 1, 0], [-1, 0, 0], [0, 0, 1]]),
        ((90.0, 90.0, 0.0), [[0, 0, -1], [-1, 0, 0], [0, 1, 0]])]


@pytest.mark.parametrize("angles,want", HAND)
def test_euler_hand_values(angles, want):
    np.testing.assert_allclose(opj.euler_angles_to_matrix(*angles), want, atol=1e-9)
    np.testing.assert_allclose(host.euler_angles_to_matrix(*angles), want, atol=1e-9)


def test_ring_index_hand_values():
    """round(sqrt((y - n//2)^2 + (x - n//2)^2)) at hand-computed pixels."""
    r8 = opj.ring_map(8)
    for (y, x), k in {(4, 4): 0, (4, 5): 1, (5, 5): 1, (6, 5): 2, (6, 6): 3, (7, 7): 4, (0, 0): 6, (1, 2): 4,
                      (4, 0): 4, (0, 4): 4, (7, 0): 5}.items():
        assert r8[y, x] == k, (y, x)
    r7 = opj.ring_map(7)  # odd size: centre 3
    for (y, x), k in {(3, 3): 0, (0, 0): 4, (6, 6): 4, (3, 6): 3, (5, 4): 2}.items():
        assert r7[y, x] == k, (y, x)


@pytest.fixture(scope="module")
def on_grid_search():
    """Noiseless, CTF-free box-64 particles generated from atoms in real space
    (independent of the projector) at grid directions, grid psi and integer
    shifts, searched by the float64 oracle with the truth as prior."""
    wl = synthetic.WORKLOADS["box64"]
    coords, elements, _ = synthetic.random_walk_chains(wl.chain_sizes, confine=wl.confine)
    numbers = np.asarray([synthetic.ELEMENT_Z.get(e, 0.0) for e in elements])
    drot, dtilt = synthetic.healpix_nest_directions(wl.healpix_order)
    rng = np.random.default_rng(11)
    B = 12
    k = rng.integers(0, len(drot), B)
    psi = wl.psi_step * rng.integers(1, 47, B)
    shifts = rng.integers(-3, 4, (B, 2))
    imgs = synthetic.whiten_batch(synthetic.project_atoms(coords, numbers, drot[k], dtilt[k], psi, shifts,
                                                          wl.box, wl.apix, wl.res)).numpy()
    vol = osp.density_from_coords(coords, numbers, wl.box, wl.apix, wl.res).astype(np.float64)
    parts = {"images": imgs, "rot": drot[k], "tilt": dtilt[k], "psi": psi, "dfu": None, "dfv": None, "dfa": None}
    recs = osr.run_search(vol, drot, dtilt, parts, apix=wl.apix, new_box=wl.box, psi_step=wl.psi_step,
                          local_range=wl.local_range, crop=10, kk=wl.kk, mask_radius=wl.mask_radius,
                          mask_edge=wl.mask_edge, dtype=np.float64)
    return dict(wl=wl, vol=vol, drot=drot, dtilt=dtilt, k=k, psi=psi, shifts=shifts, imgs=imgs, recs=recs)


def test_search_recovers_direction_and_psi(on_grid_search):
    g = on_grid_search
    recs = g["recs"]
    assert len(recs) == len(g["k"])
    hits_dir = sum((r["rot"], r["tilt"]) == (g["drot"][g["k"][r["line"]]], g["dtilt"][g["k"][r["line"]]])
                   for r in recs)
    hits_psi = sum(r["psi_fake"] == g["psi"][r["line"]] for r in recs)
    assert hits_dir >= len(recs) - 1 and hits_psi >= len(recs) - 1, (hits_dir, hits_psi)


def test_textbook_correlation_recovers_rotated_shift(on_grid_search):
    """The roll-free (textbook) correlation of the prepared template and
    particle at the true direction and psi peaks at -R(-psi) s."""
    g = on_grid_search
    wl, n = g["wl"], g["wl"].box
    mask = opj.generate_soft_mask((n, n), wl.mask_radius, wl.mask_edge)
    fsc = np.ones((n, n // 2 + 1))
    hits = 0
    for b in range(len(g["k"])):
        j = g["k"][b]
        bank = opj.build_templates(g["vol"], g["drot"][[j]], g["dtilt"][[j]], n, mask, fsc, np.float64)
        part = opt.prepare_particle(g["imgs"][b], [g["psi"][b]], n, mask, np.ones((n, n // 2 + 1)), np.float64)
        t = np.fft.ifftshift(bank[0], axes=(-2, -1))
        p = np.fft.ifftshift(part[0], axes=(-2, -1))
        ccg = np.fft.irfft2(t * p, s=(n, n))
        ly, lx = (int(v) if v < n // 2 else int(v) - n for v in np.unravel_index(np.argmax(ccg), ccg.shape))
        th = math.radians(-g["psi"][b])
        sy, sx = g["shifts"][b]
        want = (-round(math.cos(th) * sy - math.sin(th) * sx), -round(math.sin(th) * sy + math.cos(th) * sx))
        hits += (ly, lx) == want
    assert hits >= len(g["k"]) - 1, hits
