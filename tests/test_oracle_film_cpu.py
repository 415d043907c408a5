"""Pins of the film-simulation oracle (CPU only).

Each check is independent of the oracle's own code: a published algorithm
re-implemented with Python integers (SplitMix64, checked first against its
published test vector), a brute-force convolution with the mathematical
definition of reflect-101, an adaptive quadrature of the capture integral,
closed-form round trips, and the gates of the reference's own film tests
(pichromatic/src/film/{blur.rs, exposure/capture.rs, exposure/upsample.rs,
mod.rs, scan/invert.rs} test modules).
"""
import numpy as np
import pytest
from scipy import integrate

import oracle

MASK = (1 << 64) - 1
GAMMA = 0x9E3779B97F4A7C15


def _sm64_next(state):
    state = (state + GAMMA) & MASK
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return state, z ^ (z >> 31)


def _u64_to_f32(z: int) -> np.float32:
    """Round-to-nearest-even u64 -> f32 with integer arithmetic (no double rounding)."""
    n = z.bit_length()
    if n <= 24:
        return np.float32(z)
    shift = n - 24
    q, r = divmod(z, 1 << shift)
    half = 1 << (shift - 1)
    if r > half or (r == half and q & 1):
        q += 1
    return np.float32(float(q << shift))


def test_splitmix64_published_vector():
    # SplitMix64 (Steele, Lea, Flood 2014; Vigna's splitmix64.c) seeded 1234567
    s, out = 1234567, []
    for _ in range(5):
        s, z = _sm64_next(s)
        out.append(z)
    assert out == [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
                   16408922859458223821]


@pytest.mark.parametrize("seed,layer", [(1, 0), (7, 2), (0, 1)])
def test_grain_noise_is_the_splitmix64_row_stream(seed, layer):
    # grain.rs:64-82: row key seed*0xD1B54A32D192ED03 + layer*GAMMA + y, 12 draws per pixel
    w, h = 9, 4
    got = oracle.film_grain_noise(w, h, seed, layer)
    two64 = np.float32(2.0 ** 64)  # u64::MAX as f32
    for y in range(h):
        state = (seed * 0xD1B54A32D192ED03 + layer * GAMMA + y) & MASK
        for x in range(w):
            acc = np.float32(0.0)
            for _ in range(12):
                state, z = _sm64_next(state)
                acc = np.float32(acc + np.float32(_u64_to_f32(z) / two64))
            assert got[y, x] == np.float32(acc - np.float32(6.0)), (y, x)


def _reflect101(i, n):
    if n == 1:
        return 0
    p = 2 * n - 2
    i %= p
    return i if i < n else p - i


def _blur_brute(plane, sigma):
    r = max(int(np.ceil(3 * sigma)), 1)
    xs = np.arange(-r, r + 1, dtype=np.float64)
    k = np.exp(-xs * xs / (2 * sigma * sigma))
    k /= k.sum()
    h, w = plane.shape
    tmp = np.zeros((h, w))
    for x in range(w):
        tmp[:, x] = sum(k[j] * plane[:, _reflect101(x + j - r, w)] for j in range(2 * r + 1))
    out = np.zeros((h, w))
    for y in range(h):
        out[y, :] = sum(k[j] * tmp[_reflect101(y + j - r, h), :] for j in range(2 * r + 1))
    return out


@pytest.mark.parametrize("w,h,sigma", [(17, 11, 1.3), (5, 3, 2.7), (1, 6, 0.8), (23, 19, 0.5)])
def test_film_blur_matches_reflect101_brute_force(w, h, sigma):
    x = np.random.default_rng(w * h).uniform(0, 1, (h, w)).astype(np.float32)
    got = oracle.film_blur(x, sigma)
    np.testing.assert_allclose(got, _blur_brute(x.astype(np.float64), sigma), rtol=0, atol=2e-6)


def test_film_blur_reference_gates():
    # blur.rs tests: sigma ~ 0 is the identity, mass is conserved, an impulse stays symmetric
    x = np.random.default_rng(3).uniform(0, 1, (16, 16)).astype(np.float32)
    assert np.array_equal(oracle.film_blur(x, 0.0), x)
    imp = np.zeros((33, 33), np.float32)
    imp[16, 16] = 1.0
    b = oracle.film_blur(imp, 2.5)
    assert abs(float(b.sum()) - 1.0) < 1e-5
    np.testing.assert_allclose(b, b[::-1, ::-1], atol=1e-7)
    np.testing.assert_allclose(b, b.T, atol=1e-7)


@pytest.mark.parametrize("emulsion", [0, 1, 2])
def test_capture_lut_against_quadrature(emulsion):
    # capture.rs:37-106: f(phi) = 1 - E_z[exp(-k s(z)^2 phi)], s = exp(mu + sigma z), z ~ N(0,1)
    logs, fracs, (mu, sg, k) = oracle.film_capture_lut("ColorNeg200", emulsion)
    np.testing.assert_allclose(logs, np.linspace(-4.0, 8.0, 64), atol=1e-12)
    for lp, f in zip(logs[::3], fracs[::3]):
        phi = 10.0 ** lp
        g = lambda z: np.exp(-0.5 * z * z) / np.sqrt(2 * np.pi) * np.exp(-k * np.exp(2 * (mu + sg * z)) * phi)
        want = 1.0 - integrate.quad(g, -12, 12, epsabs=1e-12, limit=400)[0]
        assert abs(f - want) < 2e-4, (lp, f, want)
    # capture.rs tests: monotone, toe and shoulder limits
    assert np.all(np.diff(fracs) >= -1e-12)
    assert fracs[0] < 0.05 and fracs[-1] > 0.99
    # sampling interpolates linearly in log10(phi) between knots
    mid = 0.5 * (logs[20] + logs[21])
    assert abs(oracle.film_lut_sample("ColorNeg200", emulsion, 10.0 ** mid) - 0.5 * (fracs[20] + fracs[21])) < 1e-12
    assert oracle.film_lut_sample("ColorNeg200", emulsion, 10.0 ** logs[30]) == pytest.approx(fracs[30], abs=1e-12)


def test_upsample_reference_gates():
    # upsample.rs tests: gray round trip, non-negative, red favours long wavelengths
    g = 0.18
    s = oracle.film_upsample([g, g, g])
    assert np.all(s >= 0)
    np.testing.assert_allclose(oracle.film_spectrum_to_acescg(s), [g, g, g], atol=1e-6)
    red = oracle.film_upsample([0.8, 0.05, 0.02])
    assert red[12:16].sum() > red[0:4].sum()
    # linear in the (non-negative) basis weights
    np.testing.assert_allclose(oracle.film_upsample([0.4, 0.3, 0.2]) * 2, oracle.film_upsample([0.8, 0.6, 0.4]),
                               rtol=1e-12)


def test_normalized_dmin_orange_mask():
    # film/mod.rs test stock_dmin_drives_invert_not_image_guess
    d = oracle.film_normalized_dmin("ColorNeg200")
    assert abs(float(d.max()) - 1.0) < 1e-5
    assert d[0] > d[1] > d[2]


def _flat(w, h, rel, iso=200.0):
    g = oracle.absolute_gray(1.0, iso)
    return np.broadcast_to(np.asarray(rel, np.float32) * g, (h, w, 3)).astype(np.float32)


def test_midgray_developable_fraction_mid_scale():
    # film/mod.rs test midgray_developable_fraction_near_straight_line
    _, tr = oracle.film(_flat(4, 4, [0.185] * 3), trace=True)
    f = tr["latent"][1].mean()
    assert 0.2 < f < 0.7, f


def test_saturated_red_drives_cyan_dye():
    # film/mod.rs test saturated_red_affects_cyan_dye
    _, tr = oracle.film(_flat(4, 4, [0.6, 0.05, 0.02]), trace=True)
    assert tr["image_dye"][2].mean() > tr["image_dye"][0].mean()


def test_unexposed_scans_to_normalized_dmin():
    # film/mod.rs test unexposed_is_dmin: zero exposure scans to the film base (peak 1)
    out = oracle.film(np.zeros((8, 8, 3), np.float32))
    np.testing.assert_allclose(out.reshape(-1, 3).mean(0), oracle.film_normalized_dmin(), atol=1e-4)
    assert abs(float(out.max(axis=2).mean()) - 1.0) < 0.05


def test_neutral_positive_and_gray_ramp():
    # film/mod.rs tests neutral_stays_near_neutral_positive and gray_ramp_monotonic
    rels = [0.03, 0.06, 0.12, 0.185, 0.36, 0.7]
    ys = []
    for r in rels:
        out = oracle.film(_flat(16, 16, [r] * 3), output="PositiveLinear")
        m = out.reshape(-1, 3).mean(0)
        if r == 0.185:
            np.testing.assert_allclose(m, 0.185, atol=5e-3)
        ys.append(0.2722287168 * m[0] + 0.6740817658 * m[1] + 0.0536895174 * m[2])
    assert all(b > a for a, b in zip(ys, ys[1:])), ys


def test_film_deterministic_and_finite():
    x = _flat(8, 8, [0.185, 0.1, 0.05])
    a, b = oracle.film(x), oracle.film(x)
    assert np.array_equal(a, b) and np.isfinite(a).all()


# ---------------------------------------------------------------- all stocks
STOCKS_ALL = ["BwStub", "ColorNeg200", "Portra400", "Ektar100", "FujiPro400H", "EktachromeE100", "TriX400"]
AP1_Y = np.array([0.2722287168, 0.6740817658, 0.0536895174])


def test_all_stocks_run_and_stay_finite():
    # film/mod.rs tests all_stocks_load_and_validate and reversal_ektachrome_positive_slide
    for st in STOCKS_ALL:
        for output in ("NegativeLinear", "PositiveLinear"):
            out = oracle.film(_flat(16, 16, [0.185] * 3, iso=100.0), stock=st, output=output)
            assert np.isfinite(out).all(), (st, output)


def test_emulsion_counts_and_gammas():
    # layer lists of stock/*.rs; gamma claims of ektar_100.rs test ektar_loads_and_is_steeper_than_portra
    want_e = {"BwStub": 1, "ColorNeg200": 3, "Portra400": 6, "Ektar100": 6, "FujiPro400H": 7,
              "EktachromeE100": 3, "TriX400": 2}
    gam = {}
    for st, e in want_e.items():
        v = oracle.film_fit_nominal(st)
        assert len(v["d_max"]) == e, st
        gam[st] = v["gamma"]
    assert (gam["Ektar100"] < 0.36).any() and (gam["Ektar100"] > 0.55).any()
    assert ((gam["Portra400"] >= 0.5) & (gam["Portra400"] < 0.6)).all()


def test_bw_hd_curve_monotonic():
    # film/mod.rs test hd_curve_monotonic_patches (BwStub, NegativeLinear, box ISO 100)
    ys = []
    for y in [0.02, 0.05, 0.1, 0.185, 0.3, 0.5, 0.8, 1.2]:
        out = oracle.film(_flat(8, 8, [y] * 3, iso=100.0), stock="BwStub")
        ys.append(float((out.reshape(-1, 3) @ AP1_Y).mean()))
    assert all(b <= a + 1e-4 for a, b in zip(ys, ys[1:])), ys


def test_bw_unexposed_and_midgray_gates():
    # film/mod.rs tests unexposed_is_dmin and clipping_free_midgray (BwStub)
    out = oracle.film(np.zeros((8, 8, 3), np.float32), stock="BwStub")
    assert abs(float(out.max(axis=2).mean()) - 1.0) < 0.05
    out = oracle.film(_flat(8, 8, [0.185] * 3, iso=100.0), stock="BwStub")
    assert np.isfinite(out).all() and (out > 0).all() and (out < 1.5).all()


@pytest.mark.parametrize("stock", ["BwStub", "TriX400"])
def test_bw_scan_is_the_base_tint_at_every_exposure(stock):
    # A spectrally flat silver density gives a flat transmittance at every pixel, so every
    # scanned pixel is a scalar multiple of the film base's colour (the normalised Dmin).
    rng = np.random.default_rng(3)
    x = _flat(16, 16, [0.185] * 3, iso=400.0) * rng.uniform(0.05, 3.0, (16, 16, 1)).astype(np.float32)
    out = oracle.film(x, stock=stock).reshape(-1, 3)
    d = oracle.film_normalized_dmin(stock)
    np.testing.assert_allclose(out / out[:, :1], np.broadcast_to(d / d[0], out.shape), rtol=1e-5)


@pytest.mark.parametrize("stock", ["Portra400", "Ektar100", "FujiPro400H"])
def test_colour_negative_base_is_orange(stock):
    # the coloured couplers' mask absorbs blue most: film base R > G > B, peak normalised to 1
    d = oracle.film_normalized_dmin(stock)
    assert abs(float(d.max()) - 1.0) < 1e-5 and d[0] > d[1] > d[2]


def test_reversal_is_a_positive():
    # E-6 reversal (reduction.rs:36-41, f_eff = 1 - f): more exposure, more light through the slide
    ys = [float((oracle.film(_flat(8, 8, [r] * 3, iso=100.0), stock="EktachromeE100").reshape(-1, 3) @ AP1_Y).mean())
          for r in (0.05, 0.185, 0.5, 1.2)]
    assert all(b > a for a, b in zip(ys, ys[1:])), ys
