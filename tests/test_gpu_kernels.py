"""GPU parity of each HIP kernel against the CPU oracle at small boxes.

Tolerances: the FP32 kernels are compared with the oracle run in float64
(the exact value of the reference's algorithm); FP32 rounding of FFTs and
sums gives relative errors ~1e-6, the bar is 1e-4 relative (the stated
tolerance of the FP32 path), with integer outputs (argmax, shifts) identical.
"""
import sys

import numpy as np
import pytest

from oracle import fourier as ofr
from oracle import particles as opt
from oracle import project as opj
from oracle import search as osr
from oracle import splat as osp

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30))


def unshift(x):
    return np.fft.ifftshift(x, axes=(-2, -1))


@pytest.mark.parametrize("box,apix,res", [(32, 2.0, 4.0), (64, 1.5, 3.0)])
def test_splat_matches_oracle(hip_device, box, apix, res):
    import torch
    from gisapr_amd import density
    rng = np.random.default_rng(11)
    n = 400
    ext = box * apix * 0.6
    xyz = rng.uniform(-ext, ext, (n, 3)).astype(np.float32)
    xyz[:4] = [[0, 0, 0], [box * apix / 2 - 0.1, 0, 0], [-box * apix / 2, 1.0, 2.0], [1e3, 0, 0]]
    z = rng.choice([0.0, 1.0, 6.0, 7.0, 8.0, 16.0], n)
    say("splat", box, apix, res, n)
    got = density.density_from_coords(xyz, z, box, apix, res, device=hip_device).cpu().numpy()
    want = osp.density_from_coords(xyz, z, box, apix, res, accumulate=np.float64)
    err = np.abs(got.astype(np.float64) - want).max()
    assert err <= 2e-6 * np.abs(want).max() + 1e-7, err
    # exclusive upper bound: the last plane along each axis is never written
    assert np.all(got[-1] == 0) and np.all(got[:, -1] == 0) and np.all(got[:, :, -1] == 0)


@pytest.mark.parametrize("N", [32, 64, 96, 128, 192])
def test_volume_dft_matches_oracle(hip_device, N):
    import torch
    from gisapr_amd import kernels
    rng = np.random.default_rng(N)
    vol = rng.standard_normal((N, N, N)).astype(np.float32)
    say("volume_dft", N)
    got = kernels.volume_dft_natural(kernels.volume_dft(torch.from_numpy(vol).to(hip_device)), N).cpu().numpy()
    want = np.fft.ifftshift(opj.volume_dft(vol.astype(np.float64), np.float64), axes=(0, 1))
    assert got.shape == want.shape
    assert rel_l2(got, want) < 2e-6


def test_volume_dft_impulse_box256(hip_device):
    """Full-size check of the box-256 transform (a float64 oracle of the 512^3
    padded volume is too slow here): a single voxel at data position
    (a, b, c) has the closed-form spectrum
        D[k] = s^2 * exp(-2 pi i (kz*Pz + ky*Py + kx*Px) / 2N),
    P = (position + N/2 + N) mod 2N (pad, then fftshift), s = sinc(|f|) of
    the voxel's padded frequency, which pins every index of the three passes
    and the bricked store."""
    import torch
    from gisapr_amd import kernels
    N, L = 256, 512
    a, b, c = 37, 200, 131
    say("volume_dft impulse", N, (a, b, c))
    vol = torch.zeros((N, N, N), dtype=torch.float32, device=hip_device)
    vol[a, b, c] = 1.0
    got = kernels.volume_dft_natural(kernels.volume_dft(vol), N).cpu().numpy()
    assert got.shape == (L, L, N + 1)
    f = (np.array([a, b, c], np.float64) + N / 2 - N) / (2 * N)
    r = np.sqrt(np.sum(f * f))
    s2 = (np.sin(np.pi * r) / (np.pi * r)) ** 2
    P = (np.array([a, b, c]) + N // 2 + N) % L
    ez = np.exp(-2j * np.pi * np.arange(L) * P[0] / L)
    ey = np.exp(-2j * np.pi * np.arange(L) * P[1] / L)
    ex = np.exp(-2j * np.pi * np.arange(N + 1) * P[2] / L)
    worst = 0.0
    for z0 in range(0, L, 64):
        want = s2 * ez[z0:z0 + 64, None, None] * ey[None, :, None] * ex[None, None, :]
        worst = max(worst, float(np.abs(got[z0:z0 + 64] - want).max()))
    assert worst < 2e-6 * s2, worst


def _angles():
    return [0.0, 90.0, -150.0, 33.3, 179.9], [0.0, 90.0, 45.0, 120.7, 3.0]


@pytest.mark.parametrize("N,n", [(32, 32), (64, 64), (64, 48), (96, 96)])
def test_templates_match_oracle(hip_device, N, n):
    import torch
    from gisapr_amd import host, kernels
    rng = np.random.default_rng(100 + N + n)
    pts = rng.integers(N // 4, 3 * N // 4, (30, 3))
    vol = np.asarray(osp.density_from_coords(
        (pts[:, ::-1] - N // 2) * 1.0, np.full(len(pts), 6.0), N, 1.0, 2.0), np.float32)
    rot, tilt = _angles()
    mask = host.generate_soft_mask((N, N), N // 2 - 6, 4)
    fscv = np.linspace(1.0, 0.2, N).astype(np.float32)
    fscp = host.packed_fsc(fscv, n)
    say("templates", N, n)
    dev = hip_device
    rts = torch.from_numpy(np.ascontiguousarray(host.rotation_matrices_t(rot, tilt))).to(dev)
    dft = kernels.volume_dft(torch.from_numpy(vol).to(dev))
    bank = kernels.build_templates(dft, n, rts, torch.from_numpy(mask).to(dev),
                                   torch.from_numpy(fscp).to(dev)).cpu().numpy()
    want = opj.build_templates(vol.astype(np.float64), rot, tilt, n, mask, fscp, np.float64)
    got = ofr.pre_shift(bank)
    for d in range(len(rot)):
        e = rel_l2(got[d], want[d])
        assert e < 1e-4, (d, e)
    # unnormalised rows + norms (the scale is folded into the search) describe the same bank
    raw, norms = kernels.build_templates(dft, n, rts, torch.from_numpy(mask).to(dev),
                                         torch.from_numpy(fscp).to(dev), normalize=False)
    raw, norms = raw.cpu().numpy(), norms.cpu().numpy()
    np.testing.assert_allclose(np.linalg.norm(raw.reshape(len(rot), -1), axis=1), norms, rtol=1e-5)
    np.testing.assert_allclose(raw / norms[:, None, None], bank, rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize("N,n,M", [(64, 48, 300), (128, 128, 37)])
def test_batched_template_builds_equal_single_calls(hip_device, N, n, M):
    """C spectra in one template launch give exactly the rows of C single
    launches (bank and coefficient modes, norms included); 3 x 300 work
    items exceed the resident workgroups, so workgroups loop over several."""
    import torch
    from gisapr_amd import host, kernels
    say("batched template builds", N, n)
    dev = hip_device
    rng = np.random.default_rng(7 + N)
    C = 3
    rot, tilt = rng.uniform(-180, 180, M), rng.uniform(0, 180, M)
    rts = torch.from_numpy(np.ascontiguousarray(host.rotation_matrices_t(rot, tilt))).to(dev)
    mask = torch.from_numpy(host.generate_soft_mask((N, N), N // 2 - 6, 4)).to(dev)
    fsc = torch.from_numpy(host.packed_fsc(np.linspace(1.0, 0.3, N).astype(np.float32), n)).to(dev)
    dfts = torch.stack([kernels.volume_dft(torch.from_numpy(rng.standard_normal((N, N, N)).astype(np.float32))
                                           .to(dev)) for _ in range(C)])
    bank, norms = kernels.build_templates(dfts, n, rts, mask, fsc, normalize=False)
    assert bank.shape == (C * M, n, n // 2 + 1)
    kslot = torch.full((n, n // 2 + 1), -1, dtype=torch.int32)
    keep = torch.from_numpy(rng.permutation(n * (n // 2 + 1))[:101])
    kslot.view(-1)[keep] = torch.arange(101, dtype=torch.int32)
    kslot = kslot.to(dev)
    coef, cn = kernels.build_template_coefficients(dfts, n, rts, mask, fsc, kslot, 101, 112)
    for c in range(C):
        rows = slice(c * M, (c + 1) * M)
        b1, n1 = kernels.build_templates(dfts[c], n, rts, mask, fsc, normalize=False)
        assert torch.equal(bank[rows], b1) and torch.equal(norms[rows], n1), c
        k1, kn1 = kernels.build_template_coefficients(dfts[c], n, rts, mask, fsc, kslot, 101, 112)
        assert torch.equal(coef[rows], k1) and torch.equal(cn[rows], kn1), c


@pytest.mark.parametrize("N,n,with_ctf", [(32, 32, True), (64, 64, True), (64, 48, True), (64, 64, False),
                                          (64, 40, True)])
def test_particles_match_oracle(hip_device, N, n, with_ctf):
    import torch
    from gisapr_amd import host, kernels
    rng = np.random.default_rng(7 + N + n)
    B = 3
    imgs = rng.standard_normal((B, N, N)).astype(np.float32)
    psis = [12.0, 359.8, 200.0]
    step, rng_loc = 7.5, 15.0
    effs = [host.psi_sampling_for_record(p, step, True, rng_loc)[1] for p in psis]
    assert any(abs(a - 0.5) < 1e-9 for e in effs for a in e)  # the 0.5-degree quirk is exercised
    rotp = np.stack([host.psi_grid_params(N, e) for e in effs]).astype(np.float32)
    mask = host.generate_soft_mask((N, N), N // 2 - 4, 3)
    apix = 1.5
    new_apix = N / n * apix
    ctf = host.CTFTables(n, new_apix, 2.7, 300.0, 3.0)
    dfo = np.asarray([[15000.0, 14000.0, 30.0], [22000.0, 21000.0, 100.0], [12000.0, 12500.0, 170.0]], np.float32)
    dev = hip_device
    say("particles", N, n, with_ctf)
    spec, inv = kernels.prepare_particles(
        torch.from_numpy(imgs).to(dev), n, torch.from_numpy(rotp).to(dev), torch.from_numpy(mask).to(dev),
        torch.from_numpy(ctf.table).to(dev) if with_ctf else None,
        torch.from_numpy(dfo).to(dev) if with_ctf else None, ctf.kconst)
    got = ofr.pre_shift((spec * inv[..., None, None]).cpu().numpy())
    octf = opt.PackedCTF(n, new_apix, 2.7, 300.0, 3.0)
    for b in range(B):
        c = octf.one(*dfo[b]) if with_ctf else np.ones((n, n // 2 + 1), np.float32)
        want = opt.prepare_particle(imgs[b].astype(np.float64), effs[b], n, mask, c, np.float64)
        e = rel_l2(got[b], want)
        assert e < 1e-4, (b, e)
    # the same spectra written in fragment order (pads zeroed) and in padded rows, bit for bit
    args = (torch.from_numpy(imgs).to(dev), n, torch.from_numpy(rotp).to(dev), torch.from_numpy(mask).to(dev),
            torch.from_numpy(ctf.table).to(dev) if with_ctf else None,
            torch.from_numpy(dfo).to(dev) if with_ctf else None, ctf.kconst)
    say("particles fragment order", N, n, with_ctf)
    frag = kernels.FragmentPlanes(torch.full((B, len(effs[0]), kernels.fragment_plane(n)), float("nan"),
                                             dtype=torch.complex64, device=dev), n)
    fo, finv = kernels.prepare_particles(*args, out=frag)
    assert torch.equal(fo.data, kernels.pack_fragments(spec).data) and torch.equal(finv, inv)
    ld = kernels.padded_row_stride(n)
    po, pinv = kernels.prepare_particles(*args, out=torch.zeros((B, len(effs[0]), n, ld), dtype=torch.complex64,
                                                                 device=dev))
    assert torch.equal(po[..., :n // 2 + 1], spec) and torch.equal(pinv, inv)


@pytest.mark.parametrize("n,C", [(32, 8), (64, 10), (64, 16), (96, 16), (192, 16), (256, 16), (64, 20), (160, 32),
                                 (256, 32)])
def test_correlate_matches_oracle(hip_device, n, C):
    import torch
    from gisapr_amd import kernels
    rng = np.random.default_rng(n * 31 + C)
    m = n // 2 + 1
    M, B, P = 7, 4, 5

    def rand_spec(*shape):
        x = rng.standard_normal(shape + (n, m)) + 1j * rng.standard_normal(shape + (n, m))
        return x / np.sqrt((np.abs(x) ** 2).sum(axis=(-2, -1), keepdims=True))

    bank_pre = rand_spec(M)
    # particles correlated with a template shifted by a known lag, plus noise
    parts_pre = rand_spec(B, P) * 0.3
    local = [[0, 2, 5], [], [1], [3, 4, 5, 6]]  # particle 1 has no local direction (no record, v606:1467-1468)
    off = np.zeros(B + 1, np.int32)
    off[1:] = np.cumsum([len(l) for l in local])
    idx = np.asarray([k for l in local for k in l], np.int32)
    dev = hip_device
    say("correlate", n, C)
    c = kernels.correlate(
        torch.from_numpy(unshift(bank_pre).astype(np.complex64)).to(dev),
        torch.from_numpy(unshift(parts_pre).astype(np.complex64)).to(dev), None,
        torch.from_numpy(off).to(dev), torch.from_numpy(idx).to(dev), C)
    sc = c.scores.cpu().numpy()
    py = c.peak_y.cpu().numpy()
    px = c.peak_x.cpu().numpy()
    for b in range(B):
        if not local[b]:
            assert int(c.best[b]) == -1
            continue
        s, y, x = ofr.correlate(bank_pre[local[b]], parts_pre[b], C, np.float64)
        g = slice(off[b], off[b + 1])
        np.testing.assert_allclose(sc[g], s, rtol=1e-4, atol=1e-4 * np.abs(s).max())
        assert np.array_equal(py[g], y) and np.array_equal(px[g], x)
        assert int(c.best[b]) == int(np.argmax(s.reshape(-1)))
        np.testing.assert_allclose(float(c.mean[b]), s.mean(), rtol=1e-4, atol=1e-6)
        np.testing.assert_allclose(float(c.std[b]), s.std(), rtol=1e-4, atol=1e-6)
    # the conformation score skips the particle without a record
    pk = c.peak.cpu().numpy().astype(np.float64)
    want = sum(np.exp(pk[b] / 1024.0) for b in range(B) if local[b])
    np.testing.assert_allclose(float(kernels.score_sum(c.peak, c.best)), want, rtol=1e-12)


@pytest.mark.parametrize("name", ["box32", "box64", "box96"])
def test_end_to_end_matches_oracle(hip_device, name):
    """Splat -> bank -> particle prep -> search -> records against the float64
    oracle; box96 is the 200-particle, 192-direction CPU-runnable workload
    (radix-3 transforms of length 96 / 192)."""
    import torch
    from gisapr_amd import synthetic
    from gisapr_amd.engine import SearchEngine
    wl = synthetic.WORKLOADS[name]
    data = synthetic.make_workload(name)
    say("end-to-end", name)
    vol_model = osp.density_from_coords(data["coords"], data["numbers"], wl.box, wl.apix, wl.res)
    eng = SearchEngine(box=wl.box, apix=wl.apix, crop=wl.crop, psi_step=wl.psi_step, local_range=wl.local_range,
                       kk=wl.kk, mask_radius=wl.mask_radius, mask_edge=wl.mask_edge, fsc_values=data["fsc"],
                       dir_rot=data["dir_rot"], dir_tilt=data["dir_tilt"], device=hip_device)
    imgs = data["images"].numpy()
    dfo = np.stack([data["dfu"], data["dfv"], data["dfa"]], 1)
    ps = eng.add_particles(imgs, data["rot"], data["tilt"], data["psi"], defocus=dfo)
    vol = eng.density(data["coords"], data["numbers"], wl.res)
    np.testing.assert_allclose(vol.cpu().numpy(), vol_model, rtol=1e-5, atol=1e-5 * np.abs(vol_model).max())
    res = eng.score_volume(vol, ps)
    recs = eng.records(res, ps)
    parts = {"images": imgs, "rot": data["rot"], "tilt": data["tilt"], "psi": data["psi"],
             "dfu": data["dfu"], "dfv": data["dfv"], "dfa": data["dfa"]}
    want = osr.run_search(vol_model.astype(np.float64), data["dir_rot"], data["dir_tilt"], parts, apix=wl.apix,
                          new_box=wl.box, psi_step=wl.psi_step, local_range=wl.local_range, crop=wl.crop,
                          kk=wl.kk, mask_radius=wl.mask_radius, mask_edge=wl.mask_edge, fsc_values=data["fsc"],
                          dtype=np.float64)
    assert [r["line"] for r in recs] == [w["line"] for w in want]
    for r, w in zip(recs, want):
        assert (r["best_local"], r["best_psi"], r["xshift"], r["yshift"]) == \
               (w["best_local"], w["best_psi"], w["xshift"], w["yshift"]), (r, w)
        assert abs(r["rawcc"] - w["rawcc"]) <= 1e-4 * abs(w["rawcc"]) + 1e-3
    s_want = osr.score_sum([w["rawcc"] for w in want])
    assert abs(float(res.score.cpu()) - s_want) <= 1e-4 * s_want


def test_correlate_with_template_norms_equals_unit_bank(hip_device):
    """gis_correlate on an unnormalised bank with tinv = 1/norm equals the
    search on the unit-norm bank (scores to float32 rounding, same peaks)."""
    import torch
    from gisapr_amd import kernels
    rng = np.random.default_rng(4)
    n, C, M, B, P = 64, 10, 7, 5, 3
    m = n // 2 + 1
    d = hip_device
    raw = (rng.standard_normal((M, n, m)) + 1j * rng.standard_normal((M, n, m))).astype(np.complex64) * 37.0
    norms = np.linalg.norm(raw.reshape(M, -1), axis=1).astype(np.float32)
    unit = (raw / norms[:, None, None]).astype(np.complex64)
    spec = (rng.standard_normal((B, P, n, m)) + 1j * rng.standard_normal((B, P, n, m))).astype(np.complex64)
    off = torch.tensor([0, 2, 2, 5, 6, 7], dtype=torch.int32, device=d)
    idx = torch.tensor([0, 3, 1, 2, 6, 5, 4], dtype=torch.int32, device=d)
    inv = torch.from_numpy((1.0 / np.linalg.norm(spec.reshape(B, P, -1), axis=2)).astype(np.float32)).to(d)
    sp = torch.from_numpy(spec).to(d)
    a = kernels.correlate(torch.from_numpy(unit).to(d), sp, inv, off, idx, C)
    b = kernels.correlate(torch.from_numpy(raw).to(d), sp, inv, off, idx, C,
                          tinv=torch.from_numpy(1.0 / norms).to(d))
    np.testing.assert_allclose(b.scores.cpu().numpy(), a.scores.cpu().numpy(), rtol=2e-5, atol=1e-4)
    np.testing.assert_array_equal(b.best.cpu().numpy(), a.best.cpu().numpy())
