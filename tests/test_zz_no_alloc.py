"""The allocation-free search entry point (gis_correlate_split_noalloc) gives
the same outputs, bit for bit, as gis_correlate_split with its own workspace:
both run the same kernels, so scores, peaks, best pairs and the per-particle
statistics must be identical.  Two calls in a row, the second smaller (its
workspace, inside the library, still holds the first call's contents), then
a call above the workspace limit, which must be refused."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def make_case(torch, dev, rng, n, B, P, M):
    m = n // 2 + 1
    bank = (rng.standard_normal((M, n, m)) + 1j * rng.standard_normal((M, n, m))).astype(np.complex64)
    spectra = (rng.standard_normal((B, P, n, m)) + 1j * rng.standard_normal((B, P, n, m))).astype(np.complex64)
    counts = rng.integers(0, 6, B)
    counts[0] = 0  # a particle without local directions
    dir_idx = np.concatenate([rng.choice(M, int(c), replace=False) for c in counts]).astype(np.int32)
    dir_off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    inv_norm = (1.0 / (1.0 + rng.random((B, P)))).astype(np.float32)
    tinv = (1.0 / (1.0 + rng.random(M))).astype(np.float32)
    order = rng.permutation(B).astype(np.int32)
    to = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    return dict(bank=to(bank.view(np.float32)), spectra=to(spectra.view(np.float32)), inv_norm=to(inv_norm),
                tinv=to(tinv), dir_off=to(dir_off), dir_idx=to(dir_idx), order=to(order), B=B, P=P, n=n,
                nrows=int(dir_off[-1]))


def outputs(torch, dev, case):
    B, P, rows = case["B"], case["P"], case["nrows"]
    f32, i32, f64 = torch.float32, torch.int32, torch.float64
    # filled with garbage so that a value left unwritten cannot match by luck
    return dict(scores=torch.full((rows, P), 7.0, dtype=f32, device=dev),
                py=torch.full((rows, P), -7, dtype=i32, device=dev),
                px=torch.full((rows, P), -7, dtype=i32, device=dev),
                best=torch.full((B,), -7, dtype=i32, device=dev), peak=torch.full((B,), 7.0, dtype=f32, device=dev),
                mean=torch.full((B,), 7.0, dtype=f64, device=dev), std=torch.full((B,), 7.0, dtype=f64, device=dev))


def run(lib, case, out, crop, dt, pt, stream, ws=None):
    from gisapr_amd._lib import ptr
    args = [ptr(case["bank"]), ptr(case["spectra"]), ptr(case["inv_norm"]), ptr(case["tinv"]), ptr(case["dir_off"]),
            ptr(case["dir_idx"]), None, ptr(case["order"]), case["B"], case["P"], case["n"], crop, case["nrows"],
            case["n"] // 2 + 1, case["n"] // 2 + 1, dt, pt, ptr(out["scores"]), ptr(out["py"]), ptr(out["px"]), ptr(out["best"]), ptr(out["peak"]),
            ptr(out["mean"]), ptr(out["std"])]
    if ws is None:
        return lib.gis_correlate_split_noalloc(*args, stream)
    return lib.gis_correlate_split(*args, ptr(ws), ws.numel(), stream)


def test_noalloc_equals_workspace_entry_point(hip_device):
    import torch
    from gisapr_amd._lib import check, lib, stream_of
    L = lib()
    stream = stream_of(hip_device)
    rng = np.random.default_rng(11)
    # the second call is smaller in every dimension (box, crop, particles)
    for n, crop, B, P, M, dt, pt in [(160, 32, 37, 6, 23, 1, 0), (64, 16, 9, 3, 8, 2, 0)]:
        say("noalloc", n, crop, B, P, M, dt, pt)
        case = make_case(torch, hip_device, rng, n, B, P, M)
        wsb = int(L.gis_correlate_split_workspace(n, crop, B))
        ws = torch.empty((wsb,), dtype=torch.uint8, device=hip_device)
        want, got = outputs(torch, hip_device, case), outputs(torch, hip_device, case)
        check(run(L, case, want, crop, dt, pt, stream, ws=ws), "gis_correlate_split")
        check(run(L, case, got, crop, dt, pt, stream), "gis_correlate_split_noalloc")
        torch.cuda.synchronize(hip_device)
        assert int((want["best"] >= 0).sum()) == int((np.diff(case["dir_off"].cpu().numpy()) > 0).sum())
        for k in want:
            assert torch.equal(got[k], want[k]), (n, crop, k)


def test_noalloc_refuses_above_its_workspace(hip_device):
    """Every buffer is sized for the call (B particles without local
    directions: no search launch, only the per-particle tail), so that the
    call would stay in bounds even without the size check."""
    import torch
    from gisapr_amd._lib import lib, ptr, stream_of
    L = lib()
    n, crop, limit = 256, 32, 1 << 20
    B = 1
    while L.gis_correlate_split_workspace(n, crop, B) <= limit:
        B *= 2
    say("noalloc refuse", B)
    dev, m = hip_device, n // 2 + 1
    f32, i32, f64 = torch.float32, torch.int32, torch.float64
    bank = torch.zeros((1, n, m, 2), dtype=f32, device=dev)
    spectra = torch.zeros((1, 1, n, m, 2), dtype=f32, device=dev)
    inv_norm = torch.ones((1, 1), dtype=f32, device=dev)
    dir_off = torch.zeros((B + 1,), dtype=i32, device=dev)
    dir_idx = torch.zeros((1,), dtype=i32, device=dev)
    pmap = torch.zeros((B,), dtype=i32, device=dev)
    scores = torch.zeros((1,), dtype=f32, device=dev)
    pyx = torch.zeros((2,), dtype=i32, device=dev)
    best = torch.zeros((B,), dtype=i32, device=dev)
    peak = torch.zeros((B,), dtype=f32, device=dev)
    mean = torch.zeros((B,), dtype=f64, device=dev)
    std = torch.zeros((B,), dtype=f64, device=dev)
    rc = L.gis_correlate_split_noalloc(ptr(bank), ptr(spectra), ptr(inv_norm), None, ptr(dir_off), ptr(dir_idx),
                                       ptr(pmap), None, B, 1, n, crop, 0, m, m, 0, 0, ptr(scores), ptr(pyx),
                                       C.c_void_p(pyx.data_ptr() + 4), ptr(best), ptr(peak), ptr(mean), ptr(std),
                                       stream_of(dev))
    torch.cuda.synchronize(dev)
    assert rc != 0
    assert "limit of the allocation-free path" in L.gis_last_error().decode()
