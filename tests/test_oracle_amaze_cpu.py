"""Pins of the AMaZE restatement on textured input (CPU).

The reference's own AMaZE tests use a uniform field only, so the oracle's
textured output is pinned here by what the algorithm's structure fixes:

* Transpose equivariance.  AMaZE forms its vertical and horizontal estimates
  with the same expressions on swapped axes, and the in-place colour
  difference recurrences run along rows (hcd) and along columns (vcd) in the
  same direction, so transposing an RGGB mosaic (still RGGB) must transpose
  every plane of the passes demosaic.rs:612-786 bit for bit, with V and H
  swapped; the H/V weight of 788-828 becomes 1 - w to one rounding of its final
  division, and the Nyquist test of 830-847 (sums of four terms in another
  order) gives the same sites.
* 180-degree rotation (RGGB -> BGGR).  The pure stencil passes 612-698 agree
  to a couple of ulps (sums reordered); the in-place recurrences of 700-786 run
  the other way along each line, so they are not compared.
* The Nyquist texture test in closed form.  On a grey mosaic of vertical
  stripes of amplitude a plus a vertical ramp of slope s, every non-green site
  has hcd = +-a and vcd = 0 and every site has delh = 0, delv = 2 s; the
  normalised odd Gaussian sums to 1 and the gradient Gaussian (times
  nyqthresh = 0.5) to 0.5, so the test value is a^2 - 2 s^2: sites are Nyquist
  sites when s / a < 1 / sqrt(2).  Pure stripes trigger it and are reproduced
  exactly; a smooth ramp does not.

The full output is not equivariant: the G-at-R/B pass smooths the H/V weights
in place in raster order (demosaic.rs:949-958), the diagonal weights likewise
(1109-1123), and the chroma pass's own stencils are not mirror images of each
other (e.g. the SW estimate's last term reads dgrb at indx - M1 and its weight
mixes M3 with P3, demosaic.rs:1212-1233), so the output is only bounded.

Each pin is checked to fail on a planted mistake (oracle test hook): a
one-sided index in the horizontal gradient (plant 1), in the horizontal
colour-ratio denominator (2) and in the horizontal chain clamp (3), a bias in
the H/V weight (4), and the Nyquist threshold scaled by 10 (5) or 1/10 (6).
Tiles: 96 x 96 frames are one reference tile (top = left = -16) whose
reflect-101 border fill is symmetric except the reference's corner fill
(demosaic.rs:527-607 reads rows 32 - rr there), so comparisons stay 20 pixels
inside the tile.
"""
import numpy as np
import pytest

import oracle
from imops_amd import synth_raw_u16

PAIRS = [("dirwts0", "dirwts1"), ("delhvsqsum", "delhvsqsum"), ("vcd_pre", "hcd_pre"), ("vcdalt", "hcdalt"),
         ("dgintv", "dginth"), ("vcd", "hcd"), ("cddiffsq", "cddiffsq")]
ULP1 = float(np.spacing(np.float32(1.0)))


def _frame(seed):
    return oracle.normalize_raw(synth_raw_u16(96, 96, "RGGB", seed), 600, 16383)


def _transpose_failures(plane, plant=0):
    """Exact transpose equivariance of the passes 612-786, hvwt = 1 - w' within one
    rounding, identical Nyquist sites (tile interior)."""
    _, a = oracle.amaze_traced(plane, plant=plant)
    _, b = oracle.amaze_traced(np.ascontiguousarray(plane.T), plant=plant)
    bad = []
    for x, y in PAIRS:
        for p, q in ((x, y), (y, x)):
            if not np.array_equal(a[p], b[q].T, equal_nan=True):
                bad.append(p)
    h, ht = a["hvwt"][6:122, 6:122], b["hvwt"].T[6:122, 6:122]
    m = ~np.isnan(h)
    if np.abs(h[m] - (1.0 - ht[m])).max() > ULP1:
        bad.append("hvwt")
    n, nt = a["nyquist"], b["nyquist"].T
    if not np.array_equal(n, nt, equal_nan=True):
        bad.append("nyquist")
    return bad


def _rotation_failures(plane, plant=0):
    _, a = oracle.amaze_traced(plane, plant=plant)
    _, b = oracle.amaze_traced(np.ascontiguousarray(plane[::-1, ::-1]), "BGGR", plant=plant)
    bad = []
    sl = (slice(20, 108), slice(20, 108))
    for name in ("dirwts0", "dirwts1"):
        x, y = a[name][sl], b[name][::-1, ::-1][sl]
        if (np.abs(x - y) > 2 * np.spacing(np.maximum(np.abs(x), np.abs(y)))).any():
            bad.append(name)
    if not np.array_equal(a["delhvsqsum"][sl], b["delhvsqsum"][::-1, ::-1][sl]):
        bad.append("delhvsqsum")
    for name in ("vcd_pre", "hcd_pre", "vcdalt", "hcdalt", "dgintv", "dginth"):
        if np.abs(a[name][sl] - b[name][::-1, ::-1][sl]).max() > 4e-7:
            bad.append(name)
    return bad


def _stripe_ramp(ratio, a=0.004):
    y, x = np.mgrid[0:96, 0:96]
    return (0.4 + a * (x % 2) + ratio * a * y).astype(np.float32)


def _nyquist_failures(plant=0):
    bad = []
    inner = (slice(24, 104), slice(24, 104))
    for ratio, want in ((0.6, 1.0), (0.8, 0.0)):  # boundary s / a = 1 / sqrt(2)
        _, t = oracle.amaze_traced(_stripe_ramp(ratio), plant=plant)
        if np.nanmean(t["nyquist"][inner]) != want:
            bad.append(f"stripes+ramp {ratio}")
    y, x = np.mgrid[0:96, 0:96]
    stripes = np.where(x % 2 == 0, 0.3, 0.6).astype(np.float32)
    out, t = oracle.amaze_traced(stripes, plant=plant)
    if t["do_nyquist"][0, 0] != 1 or not np.array_equal(out[8:-8, 8:-8], np.repeat(stripes[8:-8, 8:-8, None], 3, 2)):
        bad.append("stripes")
    smooth = (0.3 + 0.3 * x / 96 + 0.1 * y / 96).astype(np.float32)
    _, t = oracle.amaze_traced(smooth, plant=plant)
    if t["do_nyquist"][0, 0] != 0:
        bad.append("smooth")
    return bad


@pytest.mark.parametrize("seed", [0, 1, 2, 3, 4, 5])
def test_amaze_passes_are_transpose_equivariant(seed):
    assert _transpose_failures(_frame(seed)) == []


@pytest.mark.parametrize("seed", [0, 3, 7])
def test_amaze_stencil_passes_are_rotation_equivariant(seed):
    assert _rotation_failures(_frame(seed)) == []


def test_amaze_nyquist_test_closed_form():
    assert _nyquist_failures() == []


def test_amaze_output_symmetry_is_bounded():
    """Whole frames (several tiles): transposing changes the output only through the
    in-place smoothing and the chroma stencils named in the module docstring.
    Bounds measured over 464^2 and 752^2 frames with margin: mean |dG| < 1e-3,
    mean |dR|, |dB| < 5e-3."""
    for n, seed in ((464, 5), (560, 9)):
        p = oracle.normalize_raw(synth_raw_u16(n, n, "RGGB", seed), 600, 16383)
        a = oracle.amaze(p)
        t = oracle.amaze(np.ascontiguousarray(p.T)).transpose(1, 0, 2)
        d = np.abs(a - t)[16:-16, 16:-16]
        assert d[..., 1].mean() < 1e-3 and d[..., 0].mean() < 5e-3 and d[..., 2].mean() < 5e-3


@pytest.mark.parametrize("plant", [1, 2, 3, 4, 5, 6])
def test_planted_mistakes_are_caught(plant):
    fails = _transpose_failures(_frame(1), plant) + _rotation_failures(_frame(1), plant) + _nyquist_failures(plant)
    assert fails, f"planted mistake {plant} passed every pin"
