"""CPU checks of the product library: it loads, exports every symbol the
public header declares, and its host-side logic (stage-list parsing, the
synthetic generator) behaves like the reference's.  No kernel launches."""
import re
from pathlib import Path

import numpy as np
import pytest

import imops_amd as im
from imops_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(imops_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    declared = _declared()
    assert len(declared) > 20
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert declared == set(_lib.SIGNATURES), "ctypes signature table out of sync with the header"


def test_parse_toml_stage_list_case_insensitive():
    cfg = im.parse_config("""
# comment
[[pipeline_modules]]
name = "demosaic"
algorithm = "markesteijn"

[[pipeline_modules]]
name = "CFACOEFFS"

[[pipeline_modules]]
name = "Exp"
ev = 0

[[pipeline_modules]]
name = "CST"
target_color_space = { value = "AcesCg", description = "x" }

[[pipeline_modules]]
name = "Gamma"
gamma = 2.4
""")
    assert cfg.pipeline_modules == ["Demosaic", "CFACoeffs", "Exp", "CST", "Gamma"]


def test_parse_json_fallback():
    cfg = im.parse_config('{"pipeline_modules": [{"name": "Demosaic", "algorithm": {"value": "amaze"}},'
                          ' {"name": "bm3d", "intensity": 0.1}, {"name": "Film", "stock": "Portra400"}]}')
    assert cfg.pipeline_modules == ["Demosaic", "BM3D", "Film"]


def test_reference_imgconfig_parses():
    # the stage list of the reference's imgconfig.toml (all names are registered modules)
    text = """
[[pipeline_modules]]
name = "Demosaic"
algorithm = "markesteijn"
[[pipeline_modules]]
name = "CFACoeffs"
[[pipeline_modules]]
name = "LumaGuidedChromaDenoise"
radius = 2
epsilon = 0.01
[[pipeline_modules]]
name = "HighlightReconstruction"
[[pipeline_modules]]
name = "Vignette"
strength = 1.0
[[pipeline_modules]]
name = "Exp"
ev = 0
[[pipeline_modules]]
name = "Contrast"
c = 1.5
[[pipeline_modules]]
name = "CST"
target_color_space = "AcesCg"
[[pipeline_modules]]
name = "LCH"
lc = 1
cc = 1
hc = 1
[[pipeline_modules]]
name = "SigmoidToneMap"
[[pipeline_modules]]
name = "CST"
target_color_space = "Srgb"
"""
    assert len(im.parse_config(text).pipeline_modules) == 11


@pytest.mark.parametrize("bad,msg", [
    ('[[pipeline_modules]]\nname = "Nope"\n', "wrong pipeline module name"),
    ('[[pipeline_modules]]\nalgorithm = "amaze"\n', "name"),
    ('[[pipeline_modules]]\nname = "Demosaic"\nalgorithm = "Amaze"\n', "Invalid Demosaic config"),
    ('[[pipeline_modules]]\nname = "Exp"\nev = "one"\n', "Invalid Exp config"),
    ("not a config", "cant decode"),
    # Film seed is a u64 (film/mod.rs:58): floats and negatives are rejected, not rounded or wrapped
    ('[[pipeline_modules]]\nname = "Film"\nseed = -3\n', "Invalid Film config"),
    ('[[pipeline_modules]]\nname = "Film"\nseed = 1.5\n', "Invalid Film config"),
    ('[[pipeline_modules]]\nname = "Film"\nseed = 1e3\n', "Invalid Film config"),
    ('{"pipeline_modules": [{"name": "Film", "seed": 18446744073709551616}]}', "cant decode"),
    # film extension fields
    ('[[pipeline_modules]]\nname = "Film"\nspectral_bins = 40\n', "spectral_bins"),
    ('[[pipeline_modules]]\nname = "Film"\ngrain = "mt19937"\n', "Invalid Film config"),
    ('[[pipeline_modules]]\nname = "Film"\ngrain_sigma = -1\n', "grain_sigma"),
    ('[[pipeline_modules]]\nname = "Film"\nhalation_radius = -3\n', "halation_radius"),
    ('[[pipeline_modules]]\nname = "Film"\nstock = "Velvia"\n', "Unknown film stock"),
])
def test_parse_errors(bad, msg):
    with pytest.raises(im.ImopsError, match=msg):
        im.parse_config(bad)


def test_film_seed_accepts_the_whole_u64_range():
    for text in ('[[pipeline_modules]]\nname = "Film"\nseed = 9007199254740993\n',
                 '{"pipeline_modules": [{"name": "Film", "seed": 18446744073709551615}]}',
                 '[[pipeline_modules]]\nname = "Film"\nseed = 0\n'):
        assert im.parse_config(text).pipeline_modules == ["Film"]


def test_synth_deterministic_and_in_range():
    a = im.synth_raw_u16(96, 64, "RGGB", 3)
    b = im.synth_raw_u16(96, 64, "RGGB", 3)
    c = im.synth_raw_u16(96, 64, "RGGB", 4)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.max() <= 16383
    s = im.synth_scene_rgb(96, 64, 3)
    assert s.min() >= 0 and s.max() <= 1
    # noise model: mean DN tracks the scene through the CFA (black 600, 15783 DN full scale)
    big = im.synth_raw_u16(512, 512, "RGGB", 1).astype(np.float64)
    sc = im.synth_scene_rgb(512, 512, 1)
    pred = 600 + sc[0::2, 0::2, 0] * 15783
    assert abs((big[0::2, 0::2] - pred).mean()) < 2.0


def test_fit_adam_matches_torch_adam():
    """The film fit's host optimiser takes torch.optim.Adam's steps (default betas and eps)
    on the same gradient sequence, including the clamp to the lower bounds after a step."""
    import torch
    from imops_amd.film_fit import Adam

    rng = np.random.default_rng(11)
    theta0 = rng.uniform(0.0, 2.0, 50).astype(np.float32)
    lower = np.full(50, 0.05, np.float32)
    ours = theta0.copy()
    opt = Adam(ours, lr=0.01)
    ref = torch.tensor(theta0.copy(), requires_grad=True)
    topt = torch.optim.Adam([ref], lr=0.01, foreach=False)
    for _ in range(200):
        g = rng.normal(size=50).astype(np.float32) * np.float32(rng.uniform(0.01, 100.0))
        opt.step(g)
        np.maximum(ours, lower, out=ours)
        ref.grad = torch.from_numpy(g.copy())
        topt.step()
        with torch.no_grad():
            ref.clamp_(min=0.05)
    np.testing.assert_allclose(ours, ref.detach().numpy(), rtol=2e-5, atol=2e-6)
    assert not np.array_equal(ours, theta0)
