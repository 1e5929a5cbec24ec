"""Pins the oracle's packed layout, pre-shift and correlation operator
(v606:494-541, 1418-1449) with independent derivations and brute force."""
import numpy as np
import pytest

from oracle import fourier as ofr


@pytest.mark.parametrize("n", [8, 16, 32, 96])
def test_packed_layout_is_rfft2(n):
    x = np.random.default_rng(n).standard_normal((n, n))
    packed = ofr.centered_full_to_packed(np.fft.fftshift(np.fft.fft2(x)))
    np.testing.assert_allclose(packed, np.fft.rfft2(x), atol=1e-10)


def test_pre_shift_rolls():
    n = 16
    a = np.arange(n * (n // 2 + 1)).reshape(n, n // 2 + 1)
    s = ofr.pre_shift(a)
    assert s[0, 0] == a[n // 2, (n // 2 + 1) - (n // 2 + 1) // 2]  # rows by n/2, cols by (n/2+1)//2
    np.testing.assert_array_equal(np.fft.ifftshift(s, axes=(-2, -1)), a)


def test_crop_indices():
    np.testing.assert_array_equal(ofr.crop_indices(16, 4), [14, 15, 0, 1])


@pytest.mark.parametrize("n,C", [(16, 4), (16, 8), (32, 10)])
def test_literal_operator_equals_unshifted_derivation(n, C):
    """irfft2 of the pre-shifted product == (-1)^ly row sign and the rolled C2R
    column weights applied to the UNshifted product (the form the HIP kernel uses)."""
    rng = np.random.default_rng(n + C)
    m = n // 2 + 1
    T = rng.standard_normal((3, n, m)) + 1j * rng.standard_normal((3, n, m))
    P = rng.standard_normal((2, n, m)) + 1j * rng.standard_normal((2, n, m))
    sc, ys, xs = ofr.correlate(ofr.pre_shift(T), ofr.pre_shift(P), C, np.float64)
    phi = ofr.correlation_functional_unshifted(n, C)
    for a in range(3):
        for b in range(2):
            ccg = np.einsum("rc,rcij->ij", T[a] * P[b], phi).real
            k = int(np.argmax(ccg))
            assert sc[a, b] == pytest.approx(ccg.ravel()[k] * 1024 * n * n, rel=1e-10)
            assert (ys[a, b], xs[a, b]) == divmod(k, C)


def test_literal_operator_equals_direct_functional():
    n, C = 16, 6
    rng = np.random.default_rng(5)
    m = n // 2 + 1
    T = rng.standard_normal((2, n, m)) + 1j * rng.standard_normal((2, n, m))
    P = rng.standard_normal((2, n, m)) + 1j * rng.standard_normal((2, n, m))
    sc, _, _ = ofr.correlate(T, P, C, np.float64)
    direct = ofr.correlate_direct(T, P, C)
    np.testing.assert_allclose(sc, direct.reshape(2, 2, -1).max(-1) * 1024 * n * n, rtol=1e-10)


def test_without_preshift_operator_is_textbook_cross_correlation():
    """Unshifted operands reduce the operator to the circular cross-correlation
    of the real images, sampled at lags [-C/2, C/2), times 1024 (scale check)."""
    n, C = 16, 6
    rng = np.random.default_rng(9)
    a = rng.standard_normal((n, n))
    b = np.roll(a, (2, -1), axis=(0, 1)) + 0.01 * rng.standard_normal((n, n))
    T = np.fft.rfft2(a)[None]
    Pc = np.conj(np.fft.rfft2(b))[None]
    sc, ys, xs = ofr.correlate(T, Pc, C, np.float64)
    # brute force: irfft2(A conj B)[d] = sum_y a(y + d) b(y)
    lags = ofr.crop_indices(n, C)
    cc = np.array([[np.sum(np.roll(a, (-dy, -dx), axis=(0, 1)) * b) for dx in lags] for dy in lags])
    k = int(np.argmax(cc))
    assert sc[0, 0] == pytest.approx(cc.ravel()[k] * 1024 * n * n, rel=1e-10)
    assert (ys[0, 0], xs[0, 0]) == divmod(k, C)
    # the peak sits at the lag that undoes the (2, -1) roll: dy = -2, dx = +1
    assert (lags[ys[0, 0]] - n if lags[ys[0, 0]] >= n // 2 else lags[ys[0, 0]]) == -2
    assert lags[xs[0, 0]] == 1


def test_first_maximum_wins_ties():
    n, C = 8, 4
    T = np.zeros((1, n, n // 2 + 1), complex)
    P = np.zeros((1, n, n // 2 + 1), complex)
    sc, ys, xs = ofr.correlate(T, P, C, np.float64)
    assert (ys[0, 0], xs[0, 0]) == (0, 0) and sc[0, 0] == 0.0


def test_precision_functional_equals_literal_operator_at_small_lags():
    """Without truncation, the precision variant's Phi over lags [-3, 3]^2 gives
    exactly the literal v606 ccg at those lags (x 1024 n^2)."""
    from oracle import precision as opr
    n, C = 16, 8
    rng = np.random.default_rng(21)
    m = n // 2 + 1
    T = rng.standard_normal((2, n, m)) + 1j * rng.standard_normal((2, n, m))
    P = rng.standard_normal((3, n, m)) + 1j * rng.standard_normal((3, n, m))
    kidx = opr.disk_indices(n, 1e9)
    assert len(kidx) == n * m
    phi = opr.phi_table(n, kidx, 3)
    s = opr.scores_f64(T.reshape(2, -1)[:, kidx], P.reshape(3, -1)[:, kidx], phi)
    lags = ofr.crop_indices(n, C)  # ccg rows/cols [n-4..n-1, 0..3]
    ly, lx = opr.lag_grid(3)
    for a in range(2):
        for b in range(3):
            prod = ofr.pre_shift(T[a] * P[b])
            ccg = np.fft.irfft2(prod, s=(n, n))
            want = np.array([ccg[y % n, x % n] for y, x in zip(ly, lx)]) * 1024 * n * n
            np.testing.assert_allclose(s[a, b], want, rtol=1e-10, atol=1e-9)
    del lags


def test_disk_indices_radius():
    from oracle import precision as opr
    n = 64
    k = opr.disk_indices(n, 10.0)
    m = n // 2 + 1
    r, c = k // m, k % m
    fy = np.where(r < n // 2, r, r - n)
    assert np.all(fy ** 2 + c ** 2 <= 100) and len(k) == sum(
        1 for y in range(-32, 32) for x in range(0, 33) if y * y + x * x <= 100)


def test_small_shift_search_particle_brute_force_and_ties():
    """Pair-major first maximum over (pair, lag), pinned by explicit loops,
    with a duplicated direction planting exact ties."""
    from oracle import precision as opr
    n = 16
    rng = np.random.default_rng(33)
    kidx = opr.disk_indices(n, 5.0)
    phi = opr.phi_table(n, kidx, 2)
    K = len(kidx)
    T = rng.standard_normal((3, K)) + 1j * rng.standard_normal((3, K))
    T[2] = T[0]                                   # direction 2 ties direction 0
    P = rng.standard_normal((2, K)) + 1j * rng.standard_normal((2, K))
    b, lag, peak, mean, std, v, _ = opr.search_particle(T, P, phi)
    best = None
    allv = []
    for j in range(3):
        for p in range(2):
            s = [1024 * sum((T[j, k] * P[p, k] * phi[k, l]).real for k in range(K)) for l in range(25)]
            m = max(s)
            allv.append(m)
            if best is None or m > best[2]:
                best = (j * 2 + p, s.index(m), m)
    assert (b, lag) == best[:2] and b < 4
    assert peak == pytest.approx(best[2], rel=1e-12)
    assert mean == pytest.approx(np.mean(allv), rel=1e-12) and std == pytest.approx(np.std(allv), rel=1e-12)


def test_small_shift_host_tables_match_oracle():
    """The product's lag table (host side of the split-bf16 search): same
    coefficient set and Phi as the oracle, hi + lo within 2^-16 relative of
    Phi, and the MFMA B-operand layout inverts to the [2K x lags] operand."""
    from gisapr_amd import host
    from oracle import precision as opr
    for n, R, half in [(32, 9.0, 3), (64, 20.5, 2), (256, 67.84, 3)]:
        t = host.SmallShiftTables(n, R, half)
        np.testing.assert_array_equal(t.kidx, opr.disk_indices(n, R))
        # kslot is the inverse of kidx on the packed plane, -1 elsewhere
        assert t.kslot.shape == (n, n // 2 + 1) and t.kslot.dtype == np.int32
        np.testing.assert_array_equal(t.kslot.reshape(-1)[t.kidx], np.arange(t.K))
        assert (t.kslot >= 0).sum() == t.K
        phi = opr.phi_table(n, t.kidx.astype(np.int64), half)
        np.testing.assert_allclose(t.phi, phi, atol=1e-12)
        tiles, ks = t.b_hi.shape[:2]
        def unlayout(b):
            return host._bf16_value(b.reshape(tiles, ks, 4, 16, 8).transpose(1, 2, 4, 0, 3).reshape(ks * 32, -1))
        B = unlayout(t.b_hi) + unlayout(t.b_lo)
        S = (2 * half + 1) ** 2
        assert np.abs(B[0:2 * t.K:2, :S] - phi.real).max() <= 2.0 ** -16 * 2
        assert np.abs(B[1:2 * t.K:2, :S] + phi.imag).max() <= 2.0 ** -16 * 2
        assert not B[2 * t.K:].any() and not B[:, S:].any()
        hi_only = unlayout(t.b_hi)
        assert np.abs(hi_only[0:2 * t.K:2, :S] - phi.real).max() > 1e-4  # lo carries real bits
    K = host.SmallShiftTables(256, 256 * 1.06 / 4, 3).K   # 4 A at box 256, 1.06 A/px
    assert 7000 < K < 7400


def test_small_shift_tables_fp16_split():
    """The fp16 operand table: the same layout, hi the nearest half to B
    (no neighbouring half is closer), lo the nearest half to B - hi, and
    hi + lo within 2^-21 of Phi (|B| <= 2; fp16 hi + lo keeps ~22 bits)."""
    from gisapr_amd import host
    from oracle import precision as opr
    with pytest.raises(ValueError):
        host.SmallShiftTables(32, 9.0, 3, "fp8")
    for n, R, half in [(32, 9.0, 3), (256, 67.84, 3)]:
        t = host.SmallShiftTables(n, R, half, "fp16")
        assert t.operand == "fp16"
        tiles, ks = t.b_hi.shape[:2]

        def unlayout(b):
            return (b.reshape(tiles, ks, 4, 16, 8).transpose(1, 2, 4, 0, 3).reshape(ks * 32, -1)
                    .view(np.float16).astype(np.float64))
        hi, lo = unlayout(t.b_hi), unlayout(t.b_lo)
        phi = opr.phi_table(n, t.kidx.astype(np.int64), half)
        S = (2 * half + 1) ** 2
        want = np.zeros_like(hi)
        want[0:2 * t.K:2, :S] = phi.real
        want[1:2 * t.K:2, :S] = -phi.imag
        assert np.abs(hi + lo - want).max() <= 2.0 ** -21
        for part, target in ((hi, want), (lo, want - hi)):
            h16 = part.astype(np.float16)
            for step in (np.inf, -np.inf):
                nb = np.nextafter(h16, np.float16(step)).astype(np.float64)
                assert np.all(np.abs(target - part) <= np.abs(target - nb))
        assert np.abs(lo).max() > 0 and not hi[2 * t.K:].any() and not hi[:, S:].any()

