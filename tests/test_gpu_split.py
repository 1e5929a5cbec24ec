"""GPU parity of the full-plane split-bf16 search (gis_correlate_split).

It evaluates the same operator as the FP32 search (every packed coefficient,
every lag of the C x C crop, v606:1418-1449), so it is checked against the
float64 oracle (``oracle.fourier.correlate``) directly: scores within 1e-4
relative (the split keeps ~16 significant bits per product, FP32
accumulation; measured errors are ~1e-6), and the oracle's peak (y, x) and
best pair exactly, except where the oracle's own best and the chosen one are
within that tolerance of each other (a near-tie, where FP32 rounding alone
could pick either).
"""
import sys

import numpy as np
import pytest

from oracle import fourier as ofr
from oracle import search as osr
from oracle import splat as osp

pytestmark = pytest.mark.gpu

RTOL = 1e-4


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def unshift(x):
    return np.fft.ifftshift(x, axes=(-2, -1))


def rand_spec(rng, n, *shape):
    m = n // 2 + 1
    x = rng.standard_normal(shape + (n, m)) + 1j * rng.standard_normal(shape + (n, m))
    return x / np.sqrt((np.abs(x) ** 2).sum(axis=(-2, -1), keepdims=True))


def direct_crop(t_pre, p_pre, C):
    """Float64 ccg crops [A, B, C, C] x 1024 n^2 (the score scale)."""
    n = t_pre.shape[-2]
    out = []
    idx = ofr.crop_indices(n, C)
    for a in range(t_pre.shape[0]):
        ccg = np.fft.irfft2(t_pre[a][None] * p_pre, s=(n, n), axes=(-2, -1))
        out.append(ccg[:, idx][:, :, idx] * 1024.0 * n * n)
    return np.stack(out)


def check_particle(sc, py, px, best, crops, C, tag):
    """sc/py/px [Mi, P] from the kernel against float64 crops [Mi, P, C, C]."""
    flat = crops.reshape(crops.shape[0], crops.shape[1], -1)
    want = flat.max(axis=2)
    scale = np.abs(want).max()
    np.testing.assert_allclose(sc, want, rtol=RTOL, atol=RTOL * scale, err_msg=tag)
    for a in range(sc.shape[0]):
        for b in range(sc.shape[1]):
            k = int(py[a, b]) * C + int(px[a, b])
            kw = int(np.argmax(flat[a, b]))
            if k != kw:  # only a near-tie may differ
                assert flat[a, b, kw] - flat[a, b, k] <= RTOL * scale, (tag, a, b, k, kw)
    bw = int(np.argmax(want.reshape(-1)))
    if best != bw:
        assert want.reshape(-1)[bw] - want.reshape(-1)[best] <= RTOL * scale, (tag, best, bw)


@pytest.mark.parametrize("n,C", [(32, 8), (40, 10), (64, 16), (96, 16), (160, 32), (192, 16), (256, 16),
                                 (256, 32)])
def test_split_matches_oracle(hip_device, n, C):
    import torch
    from gisapr_amd import kernels
    rng = np.random.default_rng(n * 17 + C)
    M, B, P = 7, 4, 5
    bank_pre = rand_spec(rng, n, M) * 3.0
    parts_pre = rand_spec(rng, n, B, P) * 0.3
    # planted correlation: particle 0 psi 2 is template 2 shifted, plus noise
    parts_pre[0, 2] += 0.2 * bank_pre[2].conj()
    local = [[0, 2, 5], [], [1], [3, 4, 5, 6]]
    off = np.zeros(B + 1, np.int32)
    off[1:] = np.cumsum([len(l) for l in local])
    idx = np.asarray([k for l in local for k in l], np.int32)
    dev = hip_device
    bank = torch.from_numpy(unshift(bank_pre).astype(np.complex64)).to(dev)
    spec = torch.from_numpy(unshift(parts_pre).astype(np.complex64)).to(dev)
    norms = np.linalg.norm(unshift(bank_pre).reshape(M, -1), axis=1)
    tinv = torch.from_numpy((1.0 / norms).astype(np.float32)).to(dev)
    inv = torch.from_numpy((1.0 / np.linalg.norm(parts_pre.reshape(B, P, -1), axis=2)).astype(np.float32)).to(dev)
    d_off, d_idx = torch.from_numpy(off).to(dev), torch.from_numpy(idx).to(dev)
    order = torch.tensor([2, 0, 3, 1], dtype=torch.int32, device=dev)
    say("split", n, C)
    c = kernels.correlate_split(bank, spec, inv, d_off, d_idx, C, tinv=tinv, order=order)
    sc, py, px = c.scores.cpu().numpy(), c.peak_y.cpu().numpy(), c.peak_x.cpu().numpy()
    unit_t = bank_pre / norms[:, None, None]
    unit_p = parts_pre / np.linalg.norm(parts_pre.reshape(B, P, -1), axis=2)[..., None, None]
    for b in range(B):
        if not local[b]:
            assert int(c.best[b]) == -1
            continue
        g = slice(off[b], off[b + 1])
        crops = direct_crop(unit_t[local[b]], unit_p[b], C)
        check_particle(sc[g], py[g], px[g], int(c.best[b]), crops, C, f"n={n} C={C} b={b}")
        want = crops.reshape(len(local[b]), P, -1).max(axis=2)
        np.testing.assert_allclose(float(c.mean[b]), want.mean(), rtol=RTOL, atol=1e-6)
        np.testing.assert_allclose(float(c.std[b]), want.std(), rtol=1e-3, atol=1e-6)
    # the planted pair is found at the lag of the planted template
    assert int(c.best[0]) == 1 * P + 2


def test_split_tiles_are_bit_identical(hip_device):
    """Every (directions, psi) per-wave tile and the processing order give
    bit-identical results; the pmap form equals the direct one."""
    import torch
    from gisapr_amd import kernels
    rng = np.random.default_rng(5)
    n, C, M, B, P = 64, 12, 9, 6, 5
    dev = hip_device
    bank = torch.from_numpy(rand_spec(rng, n, M).astype(np.complex64)).to(dev)
    spec = torch.from_numpy(rand_spec(rng, n, B, P).astype(np.complex64)).to(dev)
    local = [[0, 1, 2, 3, 4], [8], [], [2, 5, 7], [0, 1, 2, 3, 4, 5, 6, 7, 8], [6, 7]]
    off = np.zeros(B + 1, np.int32)
    off[1:] = np.cumsum([len(l) for l in local])
    idx = np.asarray([k for l in local for k in l], np.int32)
    d_off, d_idx = torch.from_numpy(off).to(dev), torch.from_numpy(idx).to(dev)
    ref = kernels.correlate_split(bank, spec, None, d_off, d_idx, C, dt=1, pt=1)
    for dt in (1, 2, 4):
        for pt in (1, 2, 3, 4, 5):
            say("split tile", dt, pt)
            got = kernels.correlate_split(bank, spec, None, d_off, d_idx, C, dt=dt, pt=pt,
                                          order=torch.tensor([4, 1, 0, 5, 3, 2], dtype=torch.int32, device=dev))
            for f in ("scores", "peak_y", "peak_x", "best", "peak", "mean", "std"):
                assert torch.equal(getattr(got, f), getattr(ref, f)), (dt, pt, f)
    # zero-padded rows (the aligned layout): the same results bit for bit
    for ld in (kernels.padded_row_stride(n), n // 2 + 1 + 5):
        say("split padded rows", ld)
        got = kernels.correlate_split(kernels.pad_rows(bank, ld), kernels.pad_rows(spec, ld), None, d_off, d_idx, C,
                                      dt=2, pt=5)
        for f in ("scores", "peak_y", "peak_x", "best", "peak", "mean", "std"):
            assert torch.equal(getattr(got, f), getattr(ref, f)), (ld, f)
    # pmap: searched particles 0..B-1 mapped onto spectra rows reversed
    rev = torch.arange(B - 1, -1, -1, dtype=torch.int32, device=dev)
    spec_rev = spec.flip(0).contiguous()
    got = kernels.correlate_split(bank, spec_rev, None, d_off, d_idx, C, pmap=rev)
    assert torch.equal(got.scores, ref.scores) and torch.equal(got.best, ref.best)


def fragment_index(n):
    """Position in fragment order of each (row, column) of a packed plane, from
    the layout's definition: tile (ks, ct) of 16 rows x 16 columns at
    (ks * nct + ct) * 256, row 16 ks + 4g + 2h + e / column 16 ct + i inside
    it at ((h * 4 + g) * 16 + i) * 2 + e; a lone last column (m % 16 == 1)
    after the tiles, one value per row."""
    m = n // 2 + 1
    nks = ((n + 15) // 16 + 1) & ~1
    lone = m % 16 == 1
    nct = m // 16 if lone else (m + 15) // 16
    r, c = np.meshgrid(np.arange(n), np.arange(m), indexing="ij")
    ks, rr = np.divmod(r, 16)
    ct, i = np.divmod(c, 16)
    g, j = np.divmod(rr, 4)
    h, e = np.divmod(j, 2)
    idx = (ks * nct + ct) * 256 + ((h * 4 + g) * 16 + i) * 2 + e
    if lone:
        idx[:, m - 1] = nks * nct * 256 + np.arange(n)
    return idx, nks * nct * 256 + (nks * 16 if lone else 0)


@pytest.mark.parametrize("n", [64, 62, 40, 256])
def test_fragment_order_is_bit_identical(hip_device, n):
    """gis_pack_fragments places every coefficient where the layout's
    definition puts it (zeros elsewhere), and both searches give the same
    results bit for bit on fragment-order operands as on row-major ones."""
    import torch
    from gisapr_amd import kernels
    rng = np.random.default_rng(n)
    C, M, B, P = min(16, n), 9, 6, 5
    dev = hip_device
    bank = torch.from_numpy(rand_spec(rng, n, M).astype(np.complex64)).to(dev)
    spec = torch.from_numpy(rand_spec(rng, n, B, P).astype(np.complex64)).to(dev)
    idx, F = fragment_index(n)
    assert kernels.fragment_plane(n) == F
    say("pack fragments", n)
    fb = kernels.pack_fragments(kernels.pad_rows(bank, kernels.padded_row_stride(n)))
    fs = kernels.pack_fragments(spec)
    want = np.zeros((M, F), np.complex64)
    want[:, idx.reshape(-1)] = bank.cpu().numpy().reshape(M, -1)
    np.testing.assert_array_equal(fb.data.cpu().numpy(), want)
    want = np.zeros((B * P, F), np.complex64)
    want[:, idx.reshape(-1)] = spec.cpu().numpy().reshape(B * P, -1)
    np.testing.assert_array_equal(fs.data.cpu().numpy().reshape(B * P, F), want)
    local = [[0, 1, 2, 3, 4], [8], [], [2, 5, 7], [0, 1, 2, 3, 4, 5, 6, 7, 8], [6, 7]]
    off = np.zeros(B + 1, np.int32)
    off[1:] = np.cumsum([len(l) for l in local])
    d_off = torch.from_numpy(off).to(dev)
    d_idx = torch.from_numpy(np.asarray([k for l in local for k in l], np.int32)).to(dev)
    fields = ("scores", "peak_y", "peak_x", "best", "peak", "mean", "std")
    for fn, dts in ((kernels.correlate_split, (1, 2, 4)), (kernels.correlate, (1, 2))):
        for dt in dts:
            ref = fn(bank, spec, None, d_off, d_idx, C, dt=dt, pt=5)
            say("fragment search", n, fn.__name__, dt)
            got = fn(fb, fs, None, d_off, d_idx, C, dt=dt, pt=5)
            for f in fields:
                assert torch.equal(getattr(got, f), getattr(ref, f)), (fn.__name__, dt, f)
    with pytest.raises(ValueError, match="both 0"):  # one operand in each layout
        kernels.correlate_split(fb, spec, None, d_off, d_idx, C)


def test_split_agrees_with_fp32_search(hip_device):
    """Same operands through the FP32 search: identical peaks and best pairs
    (both evaluate the same functional; scores within 1e-4)."""
    import torch
    from gisapr_amd import kernels
    rng = np.random.default_rng(9)
    n, C, M, B, P = 128, 16, 12, 40, 5
    dev = hip_device
    bank = torch.from_numpy(rand_spec(rng, n, M).astype(np.complex64)).to(dev)
    spec = torch.from_numpy(rand_spec(rng, n, B, P).astype(np.complex64)).to(dev)
    for b in range(0, B, 3):  # planted peaks
        spec[b, b % P] += 0.3 * bank[b % M].conj()
    local = [sorted(rng.choice(M, int(rng.integers(0, 7)), replace=False).tolist()) for _ in range(B)]
    off = np.zeros(B + 1, np.int32)
    off[1:] = np.cumsum([len(l) for l in local])
    idx = np.asarray([k for l in local for k in l], np.int32)
    d_off, d_idx = torch.from_numpy(off).to(dev), torch.from_numpy(idx).to(dev)
    a = kernels.correlate(bank, spec, None, d_off, d_idx, C)
    s = kernels.correlate_split(bank, spec, None, d_off, d_idx, C)
    sa, ss = a.scores.cpu().numpy(), s.scores.cpu().numpy()
    np.testing.assert_allclose(ss, sa, rtol=RTOL, atol=RTOL * np.abs(sa).max())
    same = (a.best.cpu().numpy() == s.best.cpu().numpy())
    for b in np.flatnonzero(~same):  # near-ties only
        g = sa[off[b]:off[b + 1]].reshape(-1)
        assert g.max() - g[int(s.best[b])] <= RTOL * np.abs(sa).max(), b


@pytest.mark.parametrize("name", ["box32", "box96"])
def test_split_end_to_end_matches_oracle(hip_device, name):
    """Splat -> bank -> particle prep -> split search -> records against the
    float64 oracle (the same records as the FP32 path)."""
    from gisapr_amd import synthetic
    from gisapr_amd.engine import SearchEngine
    wl = synthetic.WORKLOADS[name]
    data = synthetic.make_workload(name)
    say("split end-to-end", name)
    vol_model = osp.density_from_coords(data["coords"], data["numbers"], wl.box, wl.apix, wl.res)
    eng = SearchEngine(box=wl.box, apix=wl.apix, crop=wl.crop, psi_step=wl.psi_step, local_range=wl.local_range,
                       kk=wl.kk, mask_radius=wl.mask_radius, mask_edge=wl.mask_edge, fsc_values=data["fsc"],
                       dir_rot=data["dir_rot"], dir_tilt=data["dir_tilt"], device=hip_device)
    imgs = data["images"].numpy()
    dfo = np.stack([data["dfu"], data["dfv"], data["dfa"]], 1)
    ps = eng.add_particles(imgs, data["rot"], data["tilt"], data["psi"], defocus=dfo)
    vol = eng.density(data["coords"], data["numbers"], wl.res)
    bank, tinv = eng.templates(vol, normalize=False)
    res = eng.search_split(bank, ps, tinv)
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


@pytest.mark.parametrize("N,n", [(64, 64), (96, 64), (64, 40), (96, 62), (256, 256)])
def test_padded_template_bank_equals_packed(hip_device, N, n):
    """The template kernel's padded-row and fragment-order outputs
    (gis_build_templates_batched_ld, ldb > n/2+1 and ldb = 0) hold the packed
    bank bit for bit, with zeros in every pad."""
    import torch
    from gisapr_amd import host, kernels
    rng = np.random.default_rng(N + n)
    say("padded bank", N, n)
    pts = rng.integers(N // 4, 3 * N // 4, (40, 3))
    vol = torch.from_numpy(osp.density_from_coords((pts - N // 2) * 1.0, np.full(40, 6.0), N, 1.0, 2.0).astype(np.float32))
    vol = vol.to(hip_device)
    rot, tilt = rng.uniform(0, 360, 12), rng.uniform(0, 180, 12)
    rts = torch.from_numpy(host.rotation_matrices_t(rot, tilt).astype(np.float32)).to(hip_device).contiguous()
    mask = torch.from_numpy(host.generate_soft_mask((N, N), N // 3, 3).astype(np.float32)).to(hip_device)
    dft = kernels.volume_dft(vol)
    ref, rn = kernels.build_templates(dft, n, rts, mask, None, normalize=False)
    ld = kernels.padded_row_stride(n)
    out = torch.full((12, n, ld), float("nan"), dtype=torch.complex64, device=hip_device)
    got, gn = kernels.build_templates(dft, n, rts, mask, None, out=out, normalize=False)
    assert torch.equal(got[..., :n // 2 + 1], ref) and torch.equal(gn, rn)
    assert bool((got[..., n // 2 + 1:] == 0).all())
    unit = kernels.build_templates(dft, n, rts, mask, None,
                                   out=torch.empty((12, n, ld), dtype=torch.complex64, device=hip_device))
    unit_ref = kernels.build_templates(dft, n, rts, mask, None)
    assert torch.equal(unit[..., :n // 2 + 1], unit_ref)
    say("fragment-order bank", N, n)
    F = kernels.fragment_plane(n)
    frag = kernels.FragmentPlanes(torch.full((12, F), float("nan"), dtype=torch.complex64, device=hip_device), n)
    fg, fnrm = kernels.build_templates(dft, n, rts, mask, None, out=frag, normalize=False)
    assert torch.equal(fg.data, kernels.pack_fragments(ref).data) and torch.equal(fnrm, rn)  # pads zero
    assert torch.equal(fg.unpack(), ref)
    fu = kernels.build_templates(dft, n, rts, mask, None,
                                 out=kernels.FragmentPlanes(torch.empty((12, F), dtype=torch.complex64,
                                                                        device=hip_device), n))
    assert torch.equal(fu.unpack(), unit_ref)
