"""GPU parity of the fp16-operand search (gis_correlate_split_f16).

The same operator as the FP32 and split-bf16 searches (every packed
coefficient, every lag of the C x C crop, v606:1418-1449) with fp16 operands:
psi spectra stored as fp16 (unit norm x 2^7), template values rounded to
fp16 per step, packed fp16 products, FP32 accumulation.  Checked against the
float64 oracle (``oracle.fourier``: the literal irfft2 of the pre-shifted
product, cropped):

* without rescoring, every pair score within 2e-3 of the particle's largest
  float64 score (fp16 keeps 11 significant bits; the error of a pair score is
  a few 1e-4 of the peak) and the best pair the oracle's except for a
  near-tie within that tolerance;
* with the split-bf16 rescoring of each particle's best pair and near ties,
  best, peak and the best pair's lag equal the split-bf16 search's
  (gis_correlate_split) bit for bit, and so meet that search's 1e-4 bar
  against the oracle.
"""
import sys

import numpy as np
import pytest

from oracle import fourier as ofr

pytestmark = pytest.mark.gpu

RTOL16 = 2e-3   # fp16 pair scores (no rescoring), relative to the particle's largest score
RTOL = 1e-4     # rescored best pairs (split-bf16), as in the split-bf16 search's tests


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


def operands(rng, n, M, B, P, dev, local):
    """Bank and spectra (unnormalised, with 1/norms) as fragment planes, the
    fp16 spectra, the CSR direction lists and the unit-norm float64 operands."""
    import torch
    from gisapr_amd import kernels
    bank_pre = rand_spec(rng, n, M) * 3.0
    parts_pre = rand_spec(rng, n, B, P) * 0.3
    parts_pre[0, 2] += 0.2 * bank_pre[min(2, M - 1)].conj()
    off = np.zeros(B + 1, np.int32)
    off[1:] = np.cumsum([len(l) for l in local])
    idx = np.asarray([k for l in local for k in l], np.int32)
    bank = kernels.pack_fragments(torch.from_numpy(unshift(bank_pre).astype(np.complex64)).to(dev))
    spec = kernels.pack_fragments(torch.from_numpy(unshift(parts_pre).astype(np.complex64)).to(dev))
    norms = np.linalg.norm(bank_pre.reshape(M, -1), axis=1)
    pnorms = np.linalg.norm(parts_pre.reshape(B, P, -1), axis=2)
    tinv = torch.from_numpy((1.0 / norms).astype(np.float32)).to(dev)
    inv = torch.from_numpy((1.0 / pnorms).astype(np.float32)).to(dev)
    spec16 = kernels.pack_fragments_f16(spec, inv)
    d_off, d_idx = torch.from_numpy(off).to(dev), torch.from_numpy(idx).to(dev)
    unit_t = bank_pre / norms[:, None, None]
    unit_p = parts_pre / pnorms[..., None, None]
    return bank, spec, spec16, tinv, inv, d_off, d_idx, off, unit_t, unit_p


@pytest.mark.parametrize("n", [32, 40, 64, 256])
def test_pack_fragments_f16_layout(hip_device, n):
    """The fp16 planes hold fp16(x * scale * inv_norm) of every packed
    coefficient at its fp16 fragment position (float32 product, then
    round-to-nearest-even), zeros in the pads."""
    import torch
    from gisapr_amd import kernels
    rng = np.random.default_rng(n)
    x = rand_spec(rng, n, 3, 2).astype(np.complex64)
    inv = rng.uniform(0.5, 2.0, (3, 2)).astype(np.float32)
    dev = hip_device
    say("pack16", n)
    frag = kernels.pack_fragments(torch.from_numpy(x).to(dev))
    h = kernels.pack_fragments_f16(frag, torch.from_numpy(inv).to(dev), scale=128.0)
    got = h.data.to(torch.complex64).cpu().numpy()
    pos = kernels.half_fragment_positions(n)
    sc = (np.float32(128.0) * inv)[..., None, None]
    want_re = (x.real * sc).astype(np.float32).astype(np.float16).astype(np.float32)
    want_im = (x.imag * sc).astype(np.float32).astype(np.float16).astype(np.float32)
    vals = got[..., pos]
    np.testing.assert_array_equal(vals.real, want_re)
    np.testing.assert_array_equal(vals.imag, want_im)
    mask = np.ones(got.shape[-1], bool)
    mask[pos.reshape(-1)] = False
    assert not np.any(got[..., mask])


@pytest.mark.parametrize("n,C", [(32, 8), (40, 10), (64, 16), (96, 16), (160, 32), (192, 16), (256, 16),
                                 (256, 32)])
def test_split16_matches_oracle(hip_device, n, C):
    from gisapr_amd import kernels
    import torch
    rng = np.random.default_rng(n * 31 + C)
    M, B, P = 7, 4, 5
    local = [[0, 2, 5], [], [1], [3, 4, 5, 6]]
    dev = hip_device
    bank, spec, spec16, tinv, inv, d_off, d_idx, off, unit_t, unit_p = operands(rng, n, M, B, P, dev, local)
    order = torch.tensor([2, 0, 3, 1], dtype=torch.int32, device=dev)
    ref = kernels.correlate_split(bank, spec, inv, d_off, d_idx, C, tinv=tinv, order=order)
    for b16 in (None, kernels.pack_fragments_f16(bank, tinv)):
        say("split16", n, C, "fp16 bank" if b16 is not None else "float32 bank")
        raw = kernels.correlate_split_f16(bank, spec16, d_off, d_idx, C, tinv=tinv, order=order, bank16=b16)
        fine = kernels.correlate_split_f16(bank, spec16, d_off, d_idx, C, tinv=tinv, order=order, refine=spec,
                                           inv_norm=inv, bank16=b16)
        check_particles(raw, fine, ref, local, off, unit_t, unit_p, P, C, f"n={n} C={C} bank16={b16 is not None}")


def check_particles(raw, fine, ref, local, off, unit_t, unit_p, P, C, tag):
    sc = raw.scores.cpu().numpy()
    for b in range(len(local)):
        if not local[b]:
            assert int(raw.best[b]) == -1 and int(fine.best[b]) == -1
            continue
        g = slice(off[b], off[b + 1])
        crops = direct_crop(unit_t[local[b]], unit_p[b], C)
        want = crops.reshape(len(local[b]), P, -1).max(axis=2)
        scale = np.abs(want).max()
        np.testing.assert_allclose(sc[g], want, rtol=0, atol=RTOL16 * scale, err_msg=f"{tag} b={b}")
        bw, bg = int(np.argmax(want.reshape(-1))), int(raw.best[b])
        assert bg == bw or want.reshape(-1)[bw] - want.reshape(-1)[bg] <= RTOL16 * scale
        # rescored: the split-bf16 search's answer, bit for bit
        assert int(fine.best[b]) == int(ref.best[b])
        assert float(fine.peak[b]) == float(ref.peak[b])
        k = off[b] * P + int(ref.best[b])
        assert int(fine.peak_y.reshape(-1)[k]) == int(ref.peak_y.reshape(-1)[k])
        assert int(fine.peak_x.reshape(-1)[k]) == int(ref.peak_x.reshape(-1)[k])
        bf = int(fine.best[b])
        assert abs(float(fine.peak[b]) - want.reshape(-1)[bf]) <= RTOL * scale
        assert want.reshape(-1)[bw] - want.reshape(-1)[bf] <= RTOL * scale
    assert int(fine.best[0]) == 1 * P + 2


def poisoned(like):
    """Result tensors shaped like ``like``'s, filled with values no search
    writes (NaN scores, -7 indices), so an entry a launch leaves unwritten
    shows."""
    import torch
    from gisapr_amd import kernels
    f = [like.scores, like.peak_y, like.peak_x, like.best, like.peak, like.mean, like.std]
    return kernels.CorrelationResult(*[torch.full_like(t, float("nan") if t.is_floating_point() else -7) for t in f])


def test_split16_tiles_and_table_split_are_consistent(hip_device):
    """Every (directions, psi) per-wave tile gives bit-identical fp16 scores;
    the fp16-hi+lo row table agrees with the fp16-only one to fp16 accuracy;
    the rescored answers do not depend on either; the pmap form equals the
    direct one."""
    import torch
    from gisapr_amd import kernels
    rng = np.random.default_rng(11)
    n, C, M, B, P = 64, 12, 9, 6, 5
    local = [[0, 1, 2, 3, 4], [8], [], [2, 5, 7], [0, 1, 2, 3, 4, 5, 6, 7, 8], [6, 7]]
    dev = hip_device
    bank, spec, spec16, tinv, inv, d_off, d_idx, off, _, _ = operands(rng, n, M, B, P, dev, local)
    fine_ref = kernels.correlate_split_f16(bank, spec16, d_off, d_idx, C, tinv=tinv, dt=1, pt=1, refine=spec,
                                           inv_norm=inv)
    for b16 in (None, kernels.pack_fragments_f16(bank, tinv)):
        ref = kernels.correlate_split_f16(bank, spec16, d_off, d_idx, C, tinv=tinv, dt=1, pt=1, bank16=b16)
        for dt in (1, 2, 4):
            for pt in (1, 3, 5):
                say("split16 tile", dt, pt, b16 is not None)
                got = kernels.correlate_split_f16(bank, spec16, d_off, d_idx, C, tinv=tinv, dt=dt, pt=pt,
                                                  bank16=b16, out=poisoned(ref))
                assert torch.equal(got.scores, ref.scores) and torch.equal(got.best, ref.best)
                assert torch.equal(got.peak_y, ref.peak_y) and torch.equal(got.peak_x, ref.peak_x)
                fine = kernels.correlate_split_f16(bank, spec16, d_off, d_idx, C, tinv=tinv, dt=dt, pt=pt,
                                                   refine=spec, inv_norm=inv, bank16=b16, out=poisoned(ref))
                assert torch.equal(fine.best, fine_ref.best) and torch.equal(fine.peak, fine_ref.peak)
    ws = kernels.correlate_split_f16(bank, spec16, d_off, d_idx, C, tinv=tinv, wsplit=True)
    scale = float(ref.peak.abs().max())
    assert float((ws.scores - ref.scores).abs().max()) <= RTOL16 * scale
    fine_ws = kernels.correlate_split_f16(bank, spec16, d_off, d_idx, C, tinv=tinv, wsplit=True, refine=spec,
                                          inv_norm=inv)
    assert torch.equal(fine_ws.best, fine_ref.best) and torch.equal(fine_ws.peak, fine_ref.peak)
    # pmap: virtual particles mapped onto the spectra rows
    pm = torch.tensor([4, 0, 3], dtype=torch.int32, device=dev)
    loc = [local[4], local[0], local[3]]
    poff = np.zeros(4, np.int32)
    poff[1:] = np.cumsum([len(l) for l in loc])
    pidx = torch.from_numpy(np.asarray([k for l in loc for k in l], np.int32)).to(dev)
    got = kernels.correlate_split_f16(bank, spec16, torch.from_numpy(poff).to(dev), pidx, C, tinv=tinv, pmap=pm,
                                      refine=spec, inv_norm=inv)
    for v, b in enumerate([4, 0, 3]):
        assert int(got.best[v]) == int(fine_ref.best[b]) and float(got.peak[v]) == float(fine_ref.peak[b])
