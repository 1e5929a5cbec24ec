"""Pins the oracle's CTF (v606:549-692), psi sampling, bilinear rotation and
particle preparation (v606:1163-1374) with closed forms and brute force."""
import math

import numpy as np
import pytest

from oracle import fourier as ofr
from oracle import particles as opt


def ctf_closed_form(n, px, cs, kv, kk, dfu, dfv, dfa, r, c):
    """Independent float64 evaluation of the weighted CTF at packed (r, c)."""
    V = kv * 1e3
    lam = 12.2643247 / math.sqrt(V * (1 + 0.978466e-6 * V))
    fy = (r if r < n // 2 else r - n) / (n * px)
    fx = (c if c < n // 2 else -n // 2) / (n * px)
    a = math.radians(dfa)
    axx = -dfu * math.cos(a) ** 2 - dfv * math.sin(a) ** 2
    ayy = -dfu * math.sin(a) ** 2 - dfv * math.cos(a) ** 2
    axy = (dfv - dfu) * math.sin(a) * math.cos(a)
    u2 = fx * fx + fy * fy
    g = (math.pi * lam * (axx * fx * fx + 2 * axy * fx * fy + ayy * fy * fy)
         + math.pi / 2 * cs * 1e7 * lam ** 3 * u2 * u2 - math.atan(0.1 / math.sqrt(0.99)))
    v = -math.sin(g)
    s = math.sqrt(u2)
    signal = math.exp(-78.7757 * s * s - 12.9121 * s + 1.28732) / (kk + 1)
    nc = math.exp(-9.32 * s * s + 2.65 * s + 0.01908) / signal
    return v * math.sqrt(1.0 / (nc + kk * v * v))


def test_ctf_matches_closed_form():
    n, px, kk = 64, 1.5, 3.0
    ctf = opt.PackedCTF(n, px, 2.7, 300.0, kk)
    out = ctf.one(15000.0, 13000.0, 35.0)
    rng = np.random.default_rng(0)
    for r, c in zip(rng.integers(0, n, 40), rng.integers(0, n // 2 + 1, 40)):
        want = ctf_closed_form(n, px, 2.7, 300.0, kk, 15000.0, 13000.0, 35.0, r, c)
        assert abs(out[r, c] - want) <= 2e-4 * max(abs(want), 1e-3), (r, c, out[r, c], want)


def test_psi_sampling_examples():
    fake, eff = opt.psi_sampling_for_record(12.0, 7.5, True, 15.0)
    # nearest = round(12/7.5)*7.5 = 15 -> 15 - 2*7.5 + i*7.5
    assert fake == [0.0, 7.5, 15.0, 22.5, 30.0]
    assert eff == [0.5, 7.5, 15.0, 22.5, 30.0]  # |angle| < 0.5 -> 0.5
    fake, _ = opt.psi_sampling_for_record(100.0, 15.0, False, 20.0)
    assert fake == [15.0 * i for i in range(24)]
    # Python round is half-to-even: 11.25 / 7.5 = 1.5 -> 2
    assert opt.psi_sampling_for_record(11.25, 7.5, True, 7.5)[0] == [7.5, 15.0, 22.5]


def test_bilinear_zeros_brute_force():
    rng = np.random.default_rng(1)
    img = rng.standard_normal((6, 7))
    sy = np.array([0.0, 5.0, 2.3, -0.5, 5.5, -2.0, 3.9])
    sx = np.array([0.0, 6.0, 4.7, 3.0, 6.5, 1.0, -0.999])
    got = opt.bilinear_zeros(img, sy, sx)
    for y, x, g in zip(sy, sx, got):
        want = 0.0
        for iy in range(6):
            for ix in range(7):
                want += max(0, 1 - abs(y - iy)) * max(0, 1 - abs(x - ix)) * img[iy, ix]
        assert abs(g - want) < 1e-12


def test_rotation_source_is_rotation_about_centre():
    N = 16
    sy, sx = opt.rotation_source(N, 90.0)
    # 90 degrees: source(y, x) = (c + -(x - c), c + (y - c)), c = N/2
    yy, xx = np.indices((N, N))
    np.testing.assert_allclose(sy, N / 2 - (xx - N / 2), atol=1e-5)
    np.testing.assert_allclose(sx, N / 2 + (yy - N / 2), atol=1e-5)


def test_prepared_particle_properties():
    N, n = 32, 32
    rng = np.random.default_rng(2)
    img = rng.standard_normal((N, N))
    mask = np.ones((N, N), np.float32)
    ctf = np.ones((n, n // 2 + 1), np.float32)
    out = opt.prepare_particle(img, [0.5, 90.0], n, mask, ctf, np.float64)
    for p in out:
        assert abs(np.linalg.norm(p) - 1.0) < 1e-12
    # the 90-degree psi is the conjugated, normalised rfft2 of the rotated image
    rot = np.rot90(img, k=1)  # sample(y, x) = img(c - (x - c), c + (y - c)) up to the centre shift
    un = np.fft.ifftshift(out[1], axes=(-2, -1))
    sy, sx = opt.rotation_source(N, 90.0)
    r = opt.bilinear_zeros(img, sy, sx)
    spec = np.conj(np.fft.rfft2(r))
    np.testing.assert_allclose(un, spec / np.linalg.norm(spec), atol=1e-6)
    del rot
