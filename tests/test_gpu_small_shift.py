"""GPU parity of the split-bf16 small-shift search against the FP64 oracle.

The HIP path computes the same truncated functional as ``oracle.precision``
(coefficients inside the resolution disk, lags [-half, half]^2); split bf16
hi + lo products with FP32 accumulation carry ~2^-17 relative error per
product, so the bar is 1e-4 relative on every pair maximum, peak, mean and
std (scaled by the particle's largest |score|), and identical top-1
(direction, psi, lag) except where the FP64 margin between the two best
candidates is itself below that bar.
"""
import sys

import numpy as np
import pytest

from oracle import particles as opt
from oracle import precision as opr
from oracle import project as opj
from oracle import search as osr
from oracle import splat as osp

pytestmark = pytest.mark.gpu

TOL = 1e-4


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def unshift(x):
    return np.fft.ifftshift(x, axes=(-2, -1))


def margin(v):
    """FP64 gap between the best and the second-best value of a 1-D array."""
    if v.size < 2:
        return np.inf
    s = np.sort(v)
    return s[-1] - s[-2]


def compare_particles(corr, T, Pk, local, P, phi, half):
    """Check a CorrelationResult of the HIP path against opr.search_particle.
    T [rows, K], Pk [B*P, K] complex (the coefficients both sides use)."""
    scores = corr.scores.cpu().numpy()
    py = corr.peak_y.cpu().numpy()
    px = corr.peak_x.cpu().numpy()
    best = corr.best.cpu().numpy()
    peak = corr.peak.cpu().numpy()
    mean = corr.mean.cpu().numpy()
    std = corr.std.cpu().numpy()
    side = 2 * half + 1
    off = 0
    worst, near_ties, agree, used = 0.0, 0, 0, 0
    for b, idx in enumerate(local):
        m = len(idx)
        if m == 0:
            assert best[b] == -1
            continue
        want_b, want_lag, want_peak, want_mean, want_std, v, lags = opr.search_particle(
            T[idx], Pk[b * P:(b + 1) * P], phi)
        got_v = scores[off:off + m].reshape(-1).astype(np.float64)
        scale = max(np.abs(v).max(), 1e-30)
        err = np.abs(got_v - v).max() / scale
        worst = max(worst, err)
        assert err <= TOL, (b, err)
        # per pair: lag identical unless the FP64 lag margin is below the bar
        s64 = opr.scores_f64(T[idx], Pk[b * P:(b + 1) * P], phi).reshape(m * P, -1)
        got_lag = (py[off:off + m].reshape(-1) * side + px[off:off + m].reshape(-1))
        for q in range(m * P):
            if margin(s64[q]) > 2 * TOL * scale:
                assert got_lag[q] == lags[q], (b, q)
        used += 1
        agree += int(best[b] == want_b and got_lag[want_b] == want_lag)
        if margin(v) > 2 * TOL * scale:
            assert best[b] == want_b, (b, best[b], want_b)
            assert got_lag[want_b] == want_lag
        else:
            near_ties += 1
        assert abs(peak[b] - want_peak) <= TOL * scale
        assert abs(mean[b] - want_mean) <= TOL * scale
        assert abs(std[b] - want_std) <= TOL * max(want_std, 1e-30) + 1e-6 * scale
        off += m
    say(f"top-1 identical on {agree}/{used} particles; {near_ties} with an FP64 margin below 2e-4 relative")
    return worst, near_ties


def test_gather_coefficients(hip_device):
    import torch
    from gisapr_amd import kernels
    rng = np.random.default_rng(1)
    n, rows = 48, 7
    m = n // 2 + 1
    spec = (rng.standard_normal((rows, n, m)) + 1j * rng.standard_normal((rows, n, m))).astype(np.complex64)
    kidx = np.sort(rng.choice(n * m, 300, replace=False)).astype(np.int32)
    scale = rng.uniform(0.5, 2.0, rows).astype(np.float32)
    d = hip_device
    out = kernels.gather_coefficients(torch.from_numpy(spec).to(d), torch.from_numpy(kidx).to(d), 304,
                                      torch.from_numpy(scale).to(d)).cpu().numpy()
    want = spec.reshape(rows, -1)[:, kidx] * scale[:, None]
    np.testing.assert_array_equal(out[:, :300], want.astype(np.complex64))
    assert not out[:, 300:].any()
    out2 = kernels.gather_coefficients(torch.from_numpy(spec).to(d), torch.from_numpy(kidx).to(d), 304)
    np.testing.assert_array_equal(out2.cpu().numpy()[:, :300], spec.reshape(rows, -1)[:, kidx])


@pytest.mark.parametrize("operand", ["bf16", "fp16"])
@pytest.mark.parametrize("n,radius,half", [(32, 9.0, 3), (64, 20.5, 3), (96, 30.0, 2), (64, 12.0, 0)])
def test_mixed_correlate_random(hip_device, n, radius, half, operand):
    import torch
    from gisapr_amd import host, kernels
    say("mixed random", n, radius, half, operand)
    rng = np.random.default_rng(n + half)
    m = n // 2 + 1
    Mtot, B, P = 17, 9, 3
    tab = host.SmallShiftTables(n, radius, half, operand)
    K, kpad = tab.K, tab.Kpad

    def unit(*shape):
        z = rng.standard_normal(shape + (n * m,)) + 1j * rng.standard_normal(shape + (n * m,))
        return z / np.linalg.norm(z, axis=-1, keepdims=True)

    T = unit(Mtot)[:, tab.kidx]
    Pk = unit(B * P)[:, tab.kidx]
    # uneven local direction lists (rows straddle 16-row groups), one empty, one with 11 directions
    sizes = [3, 0, 1, 11, 5, 2, 7, 4, 6]
    local = [sorted(rng.choice(Mtot, s, replace=False).tolist()) for s in sizes]
    off = np.zeros(B + 1, np.int32)
    off[1:] = np.cumsum(sizes)
    idx = np.asarray([k for l in local for k in l], np.int32)
    d = hip_device

    def padded(x):
        out = np.zeros((x.shape[0], kpad), np.complex64)
        out[:, :K] = x
        return torch.from_numpy(out).to(d)

    corr = kernels.correlate_mixed(padded(T), padded(Pk), P, torch.from_numpy(off).to(d),
                                   torch.from_numpy(idx).to(d),
                                   torch.from_numpy(tab.b_hi.view(np.int16)).to(d),
                                   torch.from_numpy(tab.b_lo.view(np.int16)).to(d), half,
                                   plast=torch.from_numpy(tab.phi_last).to(d), operand=operand)
    # any processing order of the rows gives bit-identical results
    perm = torch.from_numpy(rng.permutation(len(idx)).astype(np.int32)).to(d)
    corr2 = kernels.correlate_mixed(padded(T), padded(Pk), P, torch.from_numpy(off).to(d),
                                    torch.from_numpy(idx).to(d),
                                    torch.from_numpy(tab.b_hi.view(np.int16)).to(d),
                                    torch.from_numpy(tab.b_lo.view(np.int16)).to(d), half,
                                    plast=torch.from_numpy(tab.phi_last).to(d), rows=perm, operand=operand)
    for a, b in zip(corr.__dict__.values(), corr2.__dict__.values()):
        np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())
    torch.cuda.synchronize()
    phi = opr.phi_table(n, tab.kidx.astype(np.int64), half)
    T32 = T.astype(np.complex64).astype(np.complex128)
    P32 = Pk.astype(np.complex64).astype(np.complex128)
    worst, ties = compare_particles(corr, T32, P32, local, P, phi, half)
    say("worst rel err", worst, "near ties", ties)
    if operand == "fp16":  # ~22-bit products: well inside the shared bar
        assert worst <= TOL / 5, worst


def test_mixed_correlate_rejects_bad_arguments(hip_device):
    import torch
    from gisapr_amd import host, kernels
    d = hip_device
    tab = host.SmallShiftTables(32, 9.0, 3)
    tk = torch.zeros((4, tab.Kpad), dtype=torch.complex64, device=d)
    off = torch.tensor([0, 2], dtype=torch.int32, device=d)
    idx = torch.tensor([0, 1], dtype=torch.int32, device=d)
    bh = torch.from_numpy(tab.b_hi.view(np.int16)).to(d)
    bl = torch.from_numpy(tab.b_lo.view(np.int16)).to(d)
    pk = torch.zeros((2, tab.Kpad), dtype=torch.complex64, device=d)
    with pytest.raises(ValueError):
        kernels.correlate_mixed(tk, pk, 2, off, idx, bh, bl, 3)           # 49 lags without the last-lag table
    pl = torch.from_numpy(tab.phi_last).to(d)
    with pytest.raises(ValueError):
        kernels.correlate_mixed(tk, pk, 2, off, idx, bh, bl, 4)   # 81 lags > 64
    with pytest.raises(ValueError):
        kernels.correlate_mixed(tk, pk[:1], 2, off, idx, bh, bl, 3, plast=pl)
    with pytest.raises(ValueError):
        bad = torch.tensor([0, 4], dtype=torch.int32, device=d)
        kernels.correlate_mixed(tk, pk, 2, off, bad, bh, bl, 3, plast=pl)  # direction 4 of a 4-row bank
    with pytest.raises(ValueError):
        kernels.correlate_mixed(tk, pk, 2, off, idx, bh, bl, 3, plast=pl, operand="fp8")
    with pytest.raises(ValueError, match="fp16"):  # 2^14 * 6 would overflow fp16
        kernels.correlate_mixed(tk + 3.0, pk + 2.0, 2, off, idx, bh, bl, 3, plast=pl, operand="fp16")
    # the C entry point refuses an unknown operand before touching any pointer
    from gisapr_amd._lib import lib
    rc = lib().gis_correlate_mixed_ex(None, None, None, None, None, None, 1, 1, 0, 16, None, None, None, 3, 0, 0,
                                      None, None, None, None, None, None, None, 2, None)
    assert rc < 0 and b"operand" in lib().gis_last_error()


def _oracle_inputs(data, wl, tab):
    """FP64 oracle bank and particle coefficients (v606 path, unshifted, in-disk)."""
    vol = osp.density_from_coords(data["coords"], data["numbers"], wl.box, wl.apix, wl.res).astype(np.float64)
    N = n = wl.box
    mask = opj.generate_soft_mask((N, N), wl.mask_radius, wl.mask_edge)
    fsc = np.ones(N, np.float32) if data["fsc"] is None else np.asarray(data["fsc"], np.float32)
    from oracle.fourier import centered_full_to_packed
    pf = centered_full_to_packed(opj.expand_radial_values(fsc, n))
    bank = unshift(opj.build_templates(vol, data["dir_rot"], data["dir_tilt"], n, mask, pf, np.float64))
    T = bank.reshape(bank.shape[0], -1)[:, tab.kidx]
    ctf = opt.PackedCTF(n, wl.apix, wl.cs, wl.voltage, wl.kk)
    imgs = data["images"].cpu().numpy()
    local, rows = [], []
    P = None
    for i in range(len(imgs)):
        idx = osr.local_orientations(float(data["rot"][i]), float(data["tilt"][i]), data["dir_rot"],
                                     data["dir_tilt"], wl.local_range)
        _, eff = opt.psi_sampling_for_record(float(data["psi"][i]), wl.psi_step, True, wl.local_range)
        c = ctf.one(data["dfu"][i], data["dfv"][i], data["dfa"][i])
        spec = unshift(opt.prepare_particle(imgs[i], eff, n, mask, c, np.float64))
        P = spec.shape[0]
        rows.append(spec.reshape(P, -1)[:, tab.kidx])
        local.append(idx)
    return T, np.concatenate(rows), local, P


@pytest.mark.parametrize("name,resolution", [("box32", 15.0), ("box64", 6.0), ("box96", 6.0)])
def test_small_shift_end_to_end_matches_oracle(hip_device, name, resolution):
    from gisapr_amd import synthetic
    from gisapr_amd.engine import SearchEngine
    wl = synthetic.WORKLOADS[name]
    data = synthetic.make_workload(name)
    say("small-shift end-to-end", name, resolution)
    eng = SearchEngine(box=wl.box, apix=wl.apix, crop=wl.crop, psi_step=wl.psi_step, local_range=wl.local_range,
                       kk=wl.kk, mask_radius=wl.mask_radius, mask_edge=wl.mask_edge, fsc_values=data["fsc"],
                       dir_rot=data["dir_rot"], dir_tilt=data["dir_tilt"], device=hip_device)
    tab = eng.enable_small_shift(resolution=resolution, half=3)
    assert tab.radius == pytest.approx(min(wl.box * wl.apix / resolution, wl.box / 2))
    dfo = np.stack([data["dfu"], data["dfv"], data["dfa"]], 1)
    ps = eng.add_particles(data["images"].cpu().numpy(), data["rot"], data["tilt"], data["psi"], defocus=dfo)
    eng.prepare_small_shift(ps)
    vol = eng.density(data["coords"], data["numbers"], wl.res)
    res = eng.search_small_shift(eng.templates(vol), ps)
    T, Pk, local, P = _oracle_inputs(data, wl, tab)
    assert P == ps.P and local == [list(l) for l in ps.local]
    phi = opr.phi_table(wl.box, tab.kidx.astype(np.int64), 3)
    worst, ties = compare_particles(res.corr, T, Pk, local, P, phi, 3)
    say("worst rel err", worst, "near ties", ties)
    recs = eng.records(res, ps)
    for r in recs:
        assert r["xshift"] == 3 - r["peak_x"] and r["yshift"] == 3 - r["peak_y"]
    peaks = [opr.search_particle(T[l], Pk[b * P:(b + 1) * P], phi)[2] for b, l in enumerate(local) if l]
    s_want = osr.score_sum(peaks)
    assert abs(float(res.score.cpu()) - s_want) <= TOL * s_want


@pytest.mark.parametrize("name", ["box32", "box64", "box256"])
def test_template_coefficients_equal_gathered_bank(hip_device, name):
    """The coefficient output of the template kernel equals the gather of the
    unnormalised bank scaled by 1/norm, bit for bit, with zero padding; at
    box 256 over a 96-direction subset of the 768."""
    import torch
    from gisapr_amd import kernels, synthetic
    from gisapr_amd.engine import SearchEngine
    wl = synthetic.WORKLOADS[name]
    data = synthetic.make_workload(name, n_particles=4)
    say("template coefficients", name)
    eng = SearchEngine(box=wl.box, apix=wl.apix, crop=wl.crop, psi_step=wl.psi_step, local_range=wl.local_range,
                       kk=wl.kk, mask_radius=wl.mask_radius, mask_edge=wl.mask_edge, fsc_values=data["fsc"],
                       dir_rot=data["dir_rot"], dir_tilt=data["dir_tilt"], device=hip_device)
    tab = eng.enable_small_shift(resolution=4.0 if wl.box == 256 else 8.0, half=3)
    vol = eng.density(data["coords"], data["numbers"], wl.res)
    rts = eng.rts[::8].contiguous() if wl.box == 256 else eng.rts
    dft = kernels.volume_dft(vol)
    bank, norms = kernels.build_templates(dft, eng.n, rts, eng.mask, eng.fsc_packed, normalize=False)
    want = kernels.gather_coefficients(bank, eng.small_kidx, tab.Kpad, 1.0 / norms)
    got, gnorms = kernels.build_template_coefficients(dft, eng.n, rts, eng.mask, eng.fsc_packed, eng.small_kslot,
                                                      tab.K, tab.Kpad)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(gnorms.cpu().numpy(), norms.cpu().numpy())
    g = got.cpu().numpy()
    np.testing.assert_array_equal(g, want.cpu().numpy())
    assert not g[:, tab.K:].any()
    # written into a strided slice of a larger buffer (the batched layout)
    big = torch.full((2 * rts.shape[0], tab.Kpad), 7.0 + 7.0j, dtype=torch.complex64, device=hip_device)
    kernels.build_template_coefficients(dft, eng.n, rts, eng.mask, eng.fsc_packed, eng.small_kslot, tab.K,
                                        tab.Kpad, out=big[rts.shape[0]:])
    b = big.cpu().numpy()
    np.testing.assert_array_equal(b[rts.shape[0]:], g)
    assert np.all(b[:rts.shape[0]] == 7.0 + 7.0j)


@pytest.mark.parametrize("operand", ["bf16", "fp16"])
def test_small_shift_box256_subset_precision(hip_device, operand):
    """Box 256 at 1.06 A/px, 4 A disk (K ~ 7.2k), 49 lags, seeded 256-particle
    subset: the split-bf16 path against the same functional in FP64 on the
    same FP32 coefficients.  Top-1 must agree on every particle whose FP64
    margin exceeds the error bar; near ties are counted and reported."""
    import torch
    from gisapr_amd import synthetic
    from gisapr_amd.engine import SearchEngine
    wl = synthetic.WORKLOADS["box256"]
    say("small-shift box256 subset 256", operand)
    data = synthetic.make_workload("box256", n_particles=256, device=hip_device)
    eng = SearchEngine(box=wl.box, apix=wl.apix, crop=wl.crop, psi_step=wl.psi_step, local_range=wl.local_range,
                       kk=wl.kk, mask_radius=wl.mask_radius, mask_edge=wl.mask_edge, fsc_values=data["fsc"],
                       dir_rot=data["dir_rot"], dir_tilt=data["dir_tilt"], device=hip_device)
    tab = eng.enable_small_shift(resolution=4.0, half=3, operand=operand)
    assert 7000 < tab.K < 7400
    dfo = np.stack([data["dfu"], data["dfv"], data["dfa"]], 1)
    ps = eng.add_particles(data["images"], data["rot"], data["tilt"], data["psi"], defocus=dfo)
    eng.prepare_small_shift(ps)
    vol = eng.density(data["coords"], data["numbers"], wl.res)
    bank = eng.templates(vol)
    res = eng.search_small_shift(bank, ps)
    tk = eng.template_coefficients(bank).cpu().numpy()[:, :tab.K].astype(np.complex128)
    pk = ps.coeffs.cpu().numpy()[:, :tab.K].astype(np.complex128)
    torch.cuda.synchronize()
    phi = opr.phi_table(wl.box, tab.kidx.astype(np.int64), 3)
    worst, ties = compare_particles(res.corr, tk, pk, ps.local, ps.P, phi, 3)
    say("box256 worst rel err", worst, "near ties", ties, operand)
    if operand == "fp16":
        assert worst <= TOL / 5, worst
