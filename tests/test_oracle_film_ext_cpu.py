"""CPU pins of the oracle's film extensions (spectral grid, synthetic stock,
Philox grain, constant grain sigma, halation radius / strength).

None of these exists in the reference, so each is pinned by what fixes it:
the published Philox4x32-10 known-answer vectors (Random123's kat_vectors);
the CMF fit against the reference's own 20 nm CIE table; closed forms of the
film pipeline that hold on any grid (PositiveLinear maps mid-grey to 0.185,
an unexposed frame scans to the normalised film base whose peak is 1); the
grain amplitude (a constant sigma s moves the developed density by s times a
unit-RMS field); and the synthetic stock's definition (a SplitMix64 stream,
brute-forced here with Python integers).  The reference stocks' 16-bin output
is unchanged by the grid machinery (covered by test_oracle_film_cpu.py).
"""
import numpy as np
import pytest

import oracle

M64 = (1 << 64) - 1


def _scene(w, h, seed):
    from imops_amd import synth_scene_rgb
    g = oracle.absolute_gray(0.185, 200)
    return (synth_scene_rgb(w, h, seed) * np.float32(g / 0.185)).astype(np.float32)


@pytest.mark.parametrize("ctr,key,want", [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
])
def test_philox4x32_10_known_answers(ctr, key, want):
    assert [int(x) for x in oracle.film_philox(ctr, key)] == list(want)


def test_philox_grain_noise_is_standard_normal():
    n = oracle.film_grain_noise_philox(1 << 18, seed=7, layer=1).astype(np.float64)
    assert abs(n.mean()) < 0.01 and abs(n.std() - 1.0) < 0.01
    assert abs(np.mean(n[1:] * n[:-1])) < 0.01  # neighbouring counters independent
    other = oracle.film_grain_noise_philox(1 << 18, seed=7, layer=2)
    assert abs(np.corrcoef(n, other)[0, 1]) < 0.01  # layers independent


def test_cmf_fit_matches_the_reference_table():
    c16, c81 = oracle.film_cmf(16), oracle.film_cmf(81)
    at20 = c81[:, 4:65:4]  # 400, 420, ..., 700 nm of the 81-bin grid
    assert at20.shape == (3, 16)
    assert np.abs(at20 - c16).max() < 0.015  # the fit's published accuracy class
    lam = 380 + 5 * np.arange(81)
    assert lam[c81[1].argmax()] == 555 and abs(c81[1].max() - 1.0) < 0.01


def _splitmix(state):
    while True:
        state = (state + 0x9E3779B97F4A7C15) & M64
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        yield (z ^ (z >> 31)) >> 11


@pytest.mark.parametrize("seed", [0, 7, 123456789])
def test_synthetic_stock_definition(seed):
    """Sensitivities and knots of the seeded stock from its SplitMix64 definition."""
    rng = _splitmix(seed)
    u = lambda: next(rng) / 2.0 ** 53
    lam = 380.0 + 5.0 * np.arange(81)
    for e, base in enumerate((450.0, 550.0, 650.0)):
        mu1 = base + 30.0 * (u() - 0.5)
        s1 = 25.0 + 20.0 * u()
        mu2 = mu1 + 100.0 * (u() - 0.5)
        s2 = 20.0 + 30.0 * u()
        a2 = 0.15 + 0.35 * u()
        x = [-2.5 + j + 0.3 * (u() - 0.5) for j in range(5)]
        inc = [0.0] + [0.2 + u() for _ in range(4)]
        f = 0.005 + 0.965 * np.cumsum(inc) / sum(inc)
        sens = np.exp(-0.5 * ((lam - mu1) / s1) ** 2) + a2 * np.exp(-0.5 * ((lam - mu2) / s2) ** 2)
        got = oracle.film_stock_sensitivity("Synthetic", e, 81, seed)
        np.testing.assert_allclose(got, sens, rtol=1e-14, atol=1e-300)
        kx, kf = oracle.film_stock_knots("Synthetic", e, 81, seed)
        np.testing.assert_allclose(kx, x, rtol=0, atol=1e-15)
        np.testing.assert_allclose(kf, f, rtol=1e-15)
        assert np.all(np.diff(kx) > 0) and np.all(np.diff(kf) > 0) and abs(kf[-1] - 0.97) < 1e-15 and kf[0] == 0.005


@pytest.mark.parametrize("stock,bins", [("ColorNeg200", 81), ("Synthetic", 81), ("Synthetic", 16)])
def test_positive_linear_maps_midgray_to_0185_on_any_grid(stock, bins):
    g = oracle.absolute_gray(0.185, 200)
    out = oracle.film(np.full((48, 48, 3), g, np.float32), stock=stock, output="PositiveLinear", seed=0,
                      spectral_bins=bins, stock_seed=5)
    assert np.abs(out.reshape(-1, 3).mean(0) - 0.185).max() < 2e-3


@pytest.mark.parametrize("stock,bins", [("ColorNeg200", 81), ("Synthetic", 81)])
def test_unexposed_frame_scans_to_the_normalised_base(stock, bins):
    """No light: every density is the film base (mask at 0.4 Dmax) plus the toe of the
    characteristic curve (developable fraction f(0) > 0: the crystal LUT's first knot,
    0.005 on the synthetic curve), and the scan is normalised by the base's peak
    channel, so the brightest channel reads 1 less the toe's absorption (< 2e-3)."""
    out = oracle.film(np.zeros((32, 32, 3), np.float32), stock=stock, seed=1, spectral_bins=bins, stock_seed=3,
                      grain_sigma=0.0)
    px = out.reshape(-1, 3)
    assert np.allclose(px, px[0], atol=1e-7) and 0.998 < px[0].max() <= 1.0


def test_constant_grain_sigma_moves_density_by_a_unit_rms_field():
    """grain_sigma s: D = clamp(D0 + s n / rms(n)); away from the clamps the density
    change has RMS s (n is the blurred field normalised over the whole plane)."""
    rgb = _scene(96, 80, 4)
    kw = dict(stock="Synthetic", stock_seed=2, spectral_bins=81, grain="philox", seed=9)
    _, t0 = oracle.film(rgb, trace=True, grain_sigma=0.0, **kw)
    _, t1 = oracle.film(rgb, trace=True, grain_sigma=0.01, **kw)
    d = t1["image_dye"] - t0["image_dye"]
    inner = (t0["image_dye"] > 0.05) & (t0["image_dye"] < 2.0)
    for e in range(3):
        rms = np.sqrt(np.mean(d[e][inner[e]] ** 2))
        assert abs(rms - 0.01) < 1e-3, (e, rms)


def test_halation_strength_zero_and_radius_mapping():
    """strength 0 turns the bounce off; a positive strength only adds light
    (every latent plane grows); the radius sets sigma = radius / 3."""
    rgb = _scene(96, 80, 6)
    kw = dict(stock="Synthetic", stock_seed=2, spectral_bins=81, grain_sigma=0.0, seed=1)
    _, off = oracle.film(rgb, trace=True, halation_radius=40.0, halation_strength=0.0, **kw)
    _, on = oracle.film(rgb, trace=True, halation_radius=40.0, halation_strength=0.15, **kw)
    _, wide = oracle.film(rgb, trace=True, halation_radius=60.0, halation_strength=0.15, **kw)
    assert np.all(on["latent"] >= off["latent"]) and np.any(on["latent"] > off["latent"])
    assert not np.array_equal(on["latent"], wide["latent"])


def test_generated_stock_tables_are_current():
    """oracle/oracle_stock_tables.hpp is what tests/golden/make_stock_tables.py makes from
    the reference's stock sources (checked where the reference checkout is present)."""
    import importlib.util
    from pathlib import Path
    root = Path(__file__).resolve().parent
    spec = importlib.util.spec_from_file_location("make_stock_tables", root / "golden" / "make_stock_tables.py")
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    if not gen.STOCK_DIR.is_dir():
        pytest.skip("reference stock sources not present here")
    assert gen.render() == gen.OUT.read_text()
