"""Oracle: CPU restatement of the reference algorithms (test infrastructure).

Only the tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` load this package, and only as the checker / the timed CPU
baseline.  The product library never links or calls it.

Functions and the reference code they restate:

* ``amaze``        pichromatic/src/demosaic.rs:358-1379 (f32 exact order; fp64 variant for error statements)
* ``markesteijn``  pichromatic/src/demosaic.rs:108-186; ``markesteijn3`` the three-pass extension (markesteijn3.cpp); ``preview_demosaic`` demosaic.rs:189-354 (Fast/SuperFast/SuperSuperFast)
* ``normalize_raw`` demosaic.rs:61-91 + (DN-black)/(white-black) scaling (rawler apply_scaling: parity unpinned)
* ``camera_cst``   pichromatic/src/cst.rs:12-95 (colour-crate matrices: parity unpinned)
* ``cst``, ``exp``, ``gamma``, ``cfa_coeffs``  cst.rs:6-10, exp.rs:4-11, tone_map.rs:38-49, cfa_coeffs.rs:4-12
* ``bm3d``         pichromatic/src/bm3d.rs:1-699 (f32 exact order; fp64 variant); ``chroma_bm3d`` bm3d.rs:705-759
* ``bm3d_test_image`` / ``refined_sum``  the fixture and checksum of bm3d.rs:765-789
* ``encode_u8``, ``encode_f16``, ``orient``  extern_pipeline.rs:362-395 (u8), output.rs:201-209 (f16),
  output.rs:31-70 (apply_orientation); numpy, integer/IEEE conversions
* ``box_filter``, ``luma_chroma_denoise``  pichromatic/src/chroma_denoise.rs:14-152 (f32 exact order)
* ``contrast``, ``highlight``, ``lch``, ``sigmoid``, ``baseline``, ``vignette``  contrast.rs, highlight_reconstruction.rs,
  lch.rs, tone_map.rs, modules.rs:335-371, vignette.rs (Oklab/Oklch of the colour crate: parity unpinned beyond
  the published Oklab values)
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboracle.so"

_fp = C.POINTER(C.c_float)
_sz = C.c_size_t
_lib = None

_SIGS = {
    "oracle_amaze": (C.c_int, [_fp, _sz, _sz, C.c_char_p, _sz, _sz, _fp, C.c_int]),
    "oracle_amaze_hooks": (None, [C.c_int, C.c_void_p, C.c_int, C.c_int]),
    "oracle_markesteijn": (C.c_int, [_fp, _sz, _sz, C.c_char_p, _sz, _sz, _fp]),
    "oracle_preview_demosaic": (C.c_int, [_fp, _sz, _sz, C.c_char_p, _sz, _sz, C.c_int, _fp]),
    "oracle_normalize_raw": (C.c_int, [C.c_void_p, _sz, _sz, _sz, _sz, _sz, _sz, C.c_float, C.c_float, _fp]),
    "oracle_camera_forward_matrix": (C.c_int, [_fp, _sz, _fp]),
    "oracle_camera_cst": (C.c_int, [_fp, _sz, _fp, _sz, C.c_int]),
    "oracle_cst": (C.c_int, [_fp, _sz, C.c_int, C.c_int]),
    "oracle_exp": (C.c_int, [_fp, _sz, C.c_float]),
    "oracle_gamma": (C.c_int, [_fp, _sz, C.c_float]),
    "oracle_cfa_coeffs": (C.c_int, [_fp, _sz, _fp]),
    "oracle_srgb_oetf": (C.c_float, [C.c_float]),
    "oracle_bm3d": (C.c_int, [_fp, _sz, _sz, _fp, C.c_int, C.POINTER(C.c_longlong)]),
    "oracle_group_transform_1d": (C.c_int, [C.POINTER(C.c_double), C.c_int, C.c_int, C.c_int, C.c_int]),
    "oracle_bm3d_intensity_params": (C.c_int, [C.c_float, _fp]),
    "oracle_chroma_bm3d": (C.c_int, [_fp, _sz, _sz, C.c_float]),
    "oracle_bm3d_tables": (C.c_int, [_fp, _fp, _fp]),
    "oracle_bm3d_test_image": (C.c_int, [_fp, _sz, _sz]),
    "oracle_refined_sum": (C.c_double, [_fp, _sz, C.c_int]),
    "oracle_box_filter": (C.c_int, [_fp, _fp, _sz, _sz, _sz]),
    "oracle_contrast": (C.c_int, [_fp, _sz, C.c_float]),
    "oracle_highlight": (C.c_int, [_fp, _sz, C.c_float]),
    "oracle_lch": (C.c_int, [_fp, _sz, C.c_int, C.c_float, C.c_float, C.c_float]),
    "oracle_sigmoid": (C.c_int, [_fp, _sz]),
    "oracle_baseline": (C.c_int, [_fp, _sz, C.c_float, C.c_float, C.c_float, C.c_float]),
    "oracle_vignette": (C.c_int, [_fp, _sz, _sz, C.c_void_p, _sz, C.c_float]),
    "oracle_oklch": (C.c_int, [_fp, _fp]),
    "oracle_rgc": (C.c_int, [_fp, _fp]),
    "oracle_rgc_curve": (C.c_float, [C.c_float, C.c_int]),
    "oracle_luma_chroma_denoise": (C.c_int, [_fp, _sz, _sz, _sz, C.c_float]),
}

CS_SRGB, CS_LINEAR_SRGB, CS_ACESCG, CS_XYZD65 = 0, 1, 2, 3


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
        h = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            if hasattr(h, name):
                fn = getattr(h, name)
                fn.restype = res
                fn.argtypes = args
        _lib = h
    return _lib


def _f(a):
    return a.ctypes.data_as(_fp)


def _rc(code, what):
    if code != 0:
        raise ValueError(f"oracle {what} failed ({code})")


def normalize_raw(raw: np.ndarray, black: float, white: float, crop=None) -> np.ndarray:
    raw = np.ascontiguousarray(raw, np.uint16)
    h, w = raw.shape
    cx, cy, cw, ch = crop if crop is not None else (0, 0, w, h)
    out = np.empty((ch, cw), np.float32)
    _rc(lib().oracle_normalize_raw(raw.ctypes.data, w, h, cx, cy, cw, ch, black, white, _f(out)), "normalize_raw")
    return out


def amaze(cfa_plane: np.ndarray, pattern: str = "RGGB", crop_origin=(0, 0), fp64: bool = False) -> np.ndarray:
    x = np.ascontiguousarray(cfa_plane, np.float32)
    h, w = x.shape
    out = np.zeros((h, w, 3), np.float32)
    _rc(lib().oracle_amaze(_f(x), w, h, pattern.encode(), crop_origin[0], crop_origin[1], _f(out), int(fp64)),
        "amaze")
    return out


AMAZE_TRACE = ["dirwts0", "dirwts1", "delhvsqsum", "vcd_pre", "hcd_pre", "vcdalt", "hcdalt", "dgintv", "dginth",
               "vcd", "hcd", "cddiffsq", "hvwt", "nyqutest", "nyquist", "do_nyquist"]


def amaze_traced(cfa_plane: np.ndarray, pattern: str = "RGGB", tile=(-16, -16), plant: int = 0):
    """AMaZE (FP32 order) returning (output, {plane name: 128 x 128 tile plane}) for the
    tile whose top-left corner is `tile` = (top, left); `plant` selects a deliberate
    restatement mistake (test hook, 0 = none)."""
    tr = np.full((len(AMAZE_TRACE), 128, 128), np.nan, np.float32)
    L = lib()
    L.oracle_amaze_hooks(plant, tr.ctypes.data, tile[0], tile[1])
    try:
        out = amaze(cfa_plane, pattern)
    finally:
        L.oracle_amaze_hooks(0, None, 0, 0)
    return out, dict(zip(AMAZE_TRACE, tr))


def markesteijn(cfa_plane: np.ndarray, pattern: str, crop_origin=(0, 0)) -> np.ndarray:
    x = np.ascontiguousarray(cfa_plane, np.float32)
    h, w = x.shape
    out = np.zeros((h, w, 3), np.float32)
    _rc(lib().oracle_markesteijn(_f(x), w, h, pattern.encode(), crop_origin[0], crop_origin[1], _f(out)),
        "markesteijn")
    return out


def markesteijn3(cfa_plane: np.ndarray, pattern: str, crop_origin=(0, 0), passes: int = 3,
                 fp64: bool = False, plant: int = 0) -> np.ndarray:
    """The three-pass Markesteijn X-Trans demosaic (the markesteijn3 extension);
    `plant` selects a deliberate mistake (test hook, 0 = none)."""
    x = np.ascontiguousarray(cfa_plane, np.float32)
    h, w = x.shape
    out = np.zeros((h, w, 3), np.float32)
    L = lib()
    L.oracle_markesteijn3.restype = C.c_int
    L.oracle_markesteijn3.argtypes = [_fp, _sz, _sz, C.c_char_p, _sz, _sz, C.c_int, C.c_int, _fp]
    L.oracle_markesteijn3_plant.restype = None
    L.oracle_markesteijn3_plant.argtypes = [C.c_int]
    L.oracle_markesteijn3_plant(plant)
    try:
        _rc(L.oracle_markesteijn3(_f(x), w, h, pattern.encode(), crop_origin[0], crop_origin[1], passes, int(fp64),
                                  _f(out)), "markesteijn3")
    finally:
        L.oracle_markesteijn3_plant(0)
    return out


def camera_forward_matrix(cal) -> np.ndarray:
    c = np.ascontiguousarray(cal, np.float32).ravel()
    out = np.empty(9, np.float32)
    _rc(lib().oracle_camera_forward_matrix(_f(c), c.size, _f(out)), "camera_forward_matrix")
    return out.reshape(3, 3)


def camera_cst(rgb: np.ndarray, cal, target: int) -> np.ndarray:
    out = np.array(rgb, np.float32, copy=True, order="C")
    c = np.ascontiguousarray(cal, np.float32).ravel()
    _rc(lib().oracle_camera_cst(_f(out), out.size // 3, _f(c), c.size, target), "camera_cst")
    return out


def cst(rgb: np.ndarray, src: int, dst: int) -> np.ndarray:
    out = np.array(rgb, np.float32, copy=True, order="C")
    _rc(lib().oracle_cst(_f(out), out.size // 3, src, dst), "cst")
    return out


def exp(rgb: np.ndarray, ev: float) -> np.ndarray:
    out = np.array(rgb, np.float32, copy=True, order="C")
    _rc(lib().oracle_exp(_f(out), out.size // 3, ev), "exp")
    return out


def gamma(rgb: np.ndarray, g: float) -> np.ndarray:
    out = np.array(rgb, np.float32, copy=True, order="C")
    _rc(lib().oracle_gamma(_f(out), out.size // 3, g), "gamma")
    return out


def cfa_coeffs(rgb: np.ndarray, wb) -> np.ndarray:
    out = np.array(rgb, np.float32, copy=True, order="C")
    w = np.asarray(wb, np.float32)
    _rc(lib().oracle_cfa_coeffs(_f(out), out.size // 3, _f(w)), "cfa_coeffs")
    return out


def srgb_oetf(x: float) -> float:
    return lib().oracle_srgb_oetf(x)


BM3D_FIELDS = ("intensity", "sigma", "lambda", "max_dist", "chroma_scale", "detail", "stride", "window",
               "max_group", "max_candidates", "group_transform")
GROUP_TRANSFORMS = {"hadamard": 0, "haar": 1}


def bm3d_params(intensity: float, **overrides) -> dict:
    """Bm3dParams::from_intensity (bm3d.rs:22-32) plus detail strength, as a dict;
    group_transform "hadamard" (the reference's) or "haar" (extension)."""
    v = np.empty(10, np.float32)
    lib().oracle_bm3d_intensity_params(intensity, _f(v))
    p = dict(zip(BM3D_FIELDS, [float(x) for x in v] + [0.0]))
    p.update(overrides)
    return p


def group_transform_1d(x, kind: str = "hadamard", inverse: bool = False, fp64: bool = True) -> np.ndarray:
    """One column of the group transform (bm3d.rs:595-614 for "hadamard"; the Haar option)."""
    d = np.array(x, np.float64, copy=True)
    _rc(lib().oracle_group_transform_1d(d.ctypes.data_as(C.POINTER(C.c_double)), d.size, GROUP_TRANSFORMS[kind],
                                        int(inverse), int(fp64)), "group_transform_1d")
    return d


def bm3d(rgb: np.ndarray, params: dict, fp64: bool = False):
    """BM3D of an (H, W, 3) f32 image in [0,1]; returns (output, (n_refs, n_grouped_patches))."""
    out = np.array(rgb, np.float32, copy=True, order="C")
    h, w, _ = out.shape
    gt = params.get("group_transform", 0)
    params = dict(params, group_transform=GROUP_TRANSFORMS[gt] if isinstance(gt, str) else gt)
    v = np.array([params[k] for k in BM3D_FIELDS], np.float32)
    st = (C.c_longlong * 2)()
    _rc(lib().oracle_bm3d(_f(out), w, h, _f(v), int(fp64), st), "bm3d")
    return out, (st[0], st[1])


def chroma_bm3d(rgb: np.ndarray, intensity: float) -> np.ndarray:
    out = np.array(rgb, np.float32, copy=True, order="C")
    h, w, _ = out.shape
    _rc(lib().oracle_chroma_bm3d(_f(out), w, h, intensity), "chroma_bm3d")
    return out


def box_filter(plane: np.ndarray, radius: int) -> np.ndarray:
    a = np.ascontiguousarray(plane, np.float32)
    out = np.empty_like(a)
    h, w = a.shape
    _rc(lib().oracle_box_filter(_f(a), _f(out), w, h, radius), "box_filter")
    return out


def luma_chroma_denoise(rgb: np.ndarray, radius: int, epsilon: float) -> np.ndarray:
    out = np.array(rgb, np.float32, copy=True, order="C")
    h, w, _ = out.shape
    _rc(lib().oracle_luma_chroma_denoise(_f(out), w, h, radius, epsilon), "luma_chroma_denoise")
    return out


def orient(rgb: np.ndarray, orientation: int) -> np.ndarray:
    """apply_orientation (output.rs:31-70) by its index maps: Rotate90 sends (x, y) to
    (h-1-y, x) of an h-wide frame, Rotate180 to (w-1-x, h-1-y), Rotate270 to (y, w-1-x)."""
    h, w, _ = rgb.shape
    if orientation == 0:
        return rgb.copy()
    y, x = np.mgrid[0:h, 0:w]
    if orientation == 2:
        out = np.empty_like(rgb)
        out[h - 1 - y, w - 1 - x] = rgb[y, x]
        return out
    out = np.empty((w, h, 3), rgb.dtype)  # rotated: h wide, w tall
    if orientation == 1:
        out[x, h - 1 - y] = rgb[y, x]
    else:
        out[w - 1 - x, y] = rgb[y, x]
    return out


def encode_u8(rgb: np.ndarray) -> np.ndarray:
    """(clamp(c, 0, 1) * 255) as u8: f32 multiply, truncation, NaN -> 0 (extern_pipeline.rs:381-383)."""
    a = np.asarray(rgb, np.float32)
    q = np.clip(a, np.float32(0), np.float32(1)) * np.float32(255)
    return np.where(np.isnan(q), 0, np.trunc(np.nan_to_num(q))).astype(np.uint8)


def encode_f16(rgb: np.ndarray) -> np.ndarray:
    """f16::from_f32: IEEE binary16, round to nearest even, overflow to infinity."""
    return np.asarray(rgb, np.float32).astype(np.float16)


def _rgb_op(rgb, fn, *args):
    out = np.array(rgb, np.float32, copy=True, order="C")
    _rc(fn(_f(out), out.size // 3, *args), "display stage")
    return out


def contrast(rgb, c):
    return _rgb_op(rgb, lib().oracle_contrast, c)


def highlight(rgb, clip_g):
    return _rgb_op(rgb, lib().oracle_highlight, clip_g)


def lch(rgb, cs, lc, cc, hc):
    return _rgb_op(rgb, lib().oracle_lch, cs, lc, cc, hc)


def sigmoid(rgb):
    return _rgb_op(rgb, lib().oracle_sigmoid)


def baseline(rgb, ev, shutter, f_number, iso):
    return _rgb_op(rgb, lib().oracle_baseline, ev, shutter, f_number, iso)


def vignette(rgb, opcodes: bytes, strength):
    out = np.array(rgb, np.float32, copy=True, order="C")
    h, w, _ = out.shape
    buf = C.create_string_buffer(bytes(opcodes), len(opcodes))
    _rc(lib().oracle_vignette(_f(out), w, h, C.cast(buf, C.c_void_p), len(opcodes), strength), "vignette")
    return out


def oklch(lin):
    a = np.ascontiguousarray(lin, np.float32)
    out = np.empty(3, np.float32)
    lib().oracle_oklch(_f(a), _f(out))
    return out


def rgc(rgb):
    a = np.ascontiguousarray(rgb, np.float32)
    out = np.empty(3, np.float32)
    lib().oracle_rgc(_f(a), _f(out))
    return out


def preview_demosaic(plane: np.ndarray, cfa: str, factor: int, crop=(0, 0)) -> np.ndarray:
    a = np.ascontiguousarray(plane, np.float32)
    h, w = a.shape
    out = np.empty((h // factor, w // factor, 3), np.float32)
    _rc(lib().oracle_preview_demosaic(_f(a), w, h, cfa.encode(), crop[0], crop[1], factor, _f(out)), "preview_demosaic")
    return out


def bm3d_test_image(w: int, h: int) -> np.ndarray:
    out = np.empty((h, w, 3), np.float32)
    _rc(lib().oracle_bm3d_test_image(_f(out), w, h), "bm3d_test_image")
    return out


def refined_sum(rgb: np.ndarray, channel: int) -> float:
    a = np.ascontiguousarray(rgb, np.float32)
    return lib().oracle_refined_sum(_f(a), a.size // 3, channel)


STOCKS = {"BwStub": 0, "ColorNeg200": 1, "Portra400": 2, "Ektar100": 3, "FujiPro400H": 4, "EktachromeE100": 5,
          "TriX400": 6, "Synthetic": 7}
FORMATS = {"Film35mm": 0, "Film6x6": 1, "Film4x5": 2}
_dp = C.POINTER(C.c_double)



def film(rgb: np.ndarray, stock: str = "ColorNeg200", film_format: str = "Film35mm", seed: int = 1,
         output: str = "NegativeLinear", shutter: float = 0.0, fp64: bool = False, n_emulsions: int = 3,
         trace: bool = False, spectral_bins: int = 16, stock_seed: int = 0, grain: str = "splitmix",
         grain_sigma: float = -1.0, halation_radius: float = -1.0, halation_strength: float = -1.0):
    """film::process (film/mod.rs:77-129) on an (H, W, 3) absolute-luminance ACEScg image.

    Extension keywords as the film stage's (spectral_bins 16 | 81, stock "Synthetic" with
    stock_seed, grain "splitmix" | "philox", grain_sigma, halation_radius,
    halation_strength; negative: the reference's behaviour).
    Returns the scanned ACEScg image, or (image, trace) with the latent, image-dye and
    mask-dye planes (n_emulsions x H x W, f64) and the PositiveLinear mid/Dmin."""
    L = lib()
    L.oracle_film.restype = C.c_int
    L.oracle_film.argtypes = [_fp, _sz, _sz, C.c_int, C.c_int, C.c_ulonglong, C.c_int, C.c_float, C.c_int,
                              _dp, _dp, _dp, _fp, _dp]
    ext = np.array([spectral_bins, float(stock_seed), 1.0 if grain == "philox" else 0.0, grain_sigma,
                    halation_radius, halation_strength], np.float64)
    if stock_seed >= 2 ** 53:
        raise ValueError("oracle.film: stock_seed must be below 2^53")
    out = np.array(rgb, np.float32, copy=True, order="C")
    h, w, _ = out.shape
    planes = [np.zeros((n_emulsions, h, w), np.float64) for _ in range(3)] if trace else [None] * 3
    md = np.zeros(6, np.float32)
    ptr = lambda a: a.ctypes.data_as(_dp) if a is not None else None
    _rc(L.oracle_film(_f(out), w, h, STOCKS[stock], FORMATS[film_format], seed,
                      1 if output == "PositiveLinear" else 0, shutter, int(fp64), ptr(planes[0]), ptr(planes[1]),
                      ptr(planes[2]), _f(md) if trace else None, ext.ctypes.data_as(_dp)), "film")
    if not trace:
        return out
    return out, {"latent": planes[0], "image_dye": planes[1], "mask_dye": planes[2], "mid": md[:3].copy(),
                 "dmin": md[3:].copy()}


def absolute_gray(y_rel: float, box_iso: float) -> np.float32:
    """relative_to_absolute_y under sunny-16 (film/mod.rs:153-163)."""
    s = np.float32(max(box_iso, 1.0))
    t = float(np.float32(1.0) / s)
    return np.float32(12.5 * float(np.float32(y_rel)) * 64.0 / (t * float(s)))


def _film_sigs(L):
    L.oracle_film_blur.restype = None
    L.oracle_film_blur.argtypes = [_fp, _sz, _sz, C.c_float]
    L.oracle_film_grain_noise.restype = None
    L.oracle_film_grain_noise.argtypes = [_fp, _sz, _sz, C.c_ulonglong, _sz]
    L.oracle_film_capture_lut.restype = C.c_int
    L.oracle_film_capture_lut.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
    L.oracle_film_lut_sample.restype = C.c_double
    L.oracle_film_lut_sample.argtypes = [C.c_int, C.c_int, C.c_double]
    L.oracle_film_upsample.restype = None
    L.oracle_film_upsample.argtypes = [_fp, _dp]
    L.oracle_film_spectrum_to_acescg.restype = None
    L.oracle_film_spectrum_to_acescg.argtypes = [_dp, _dp]
    L.oracle_film_normalized_dmin.restype = C.c_int
    L.oracle_film_normalized_dmin.argtypes = [C.c_int, _fp]
    return L


def film_blur(plane: np.ndarray, sigma: float) -> np.ndarray:
    """gaussian_blur_separable (film/blur.rs:10-89) on an f32 (H, W) plane."""
    out = np.array(plane, np.float32, copy=True, order="C")
    _film_sigs(lib()).oracle_film_blur(_f(out), out.shape[1], out.shape[0], sigma)
    return out


def film_grain_noise(w: int, h: int, seed: int, layer: int) -> np.ndarray:
    """Unit grain noise field before its blur (grain.rs:20-45, 64-82)."""
    out = np.zeros((h, w), np.float32)
    _film_sigs(lib()).oracle_film_grain_noise(_f(out), w, h, seed, layer)
    return out


def film_philox(counter, key) -> np.ndarray:
    """Philox4x32-10 of a 4-word counter under a 2-word key (the grain extension's generator)."""
    c = np.ascontiguousarray(counter, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    L = lib()
    L.oracle_film_philox.restype = None
    L.oracle_film_philox.argtypes = [C.c_void_p] * 3
    L.oracle_film_philox(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    return out


def film_grain_noise_philox(n: int, seed: int, layer: int) -> np.ndarray:
    """Philox grain noise (before its blur) of n pixels."""
    out = np.zeros(n, np.float32)
    L = lib()
    L.oracle_film_grain_noise_philox.restype = None
    L.oracle_film_grain_noise_philox.argtypes = [_fp, _sz, C.c_ulonglong, _sz]
    L.oracle_film_grain_noise_philox(_f(out), n, seed, layer)
    return out


def film_stock_knots(stock: str, emulsion: int, spectral_bins: int = 16, stock_seed: int = 0):
    """(log10 fluence knots, developable fractions) of an emulsion's characteristic curve."""
    logs, fracs = np.zeros(64), np.zeros(64)
    L = lib()
    L.oracle_film_stock_knots.restype = C.c_int
    L.oracle_film_stock_knots.argtypes = [C.c_int, C.c_int, C.c_ulonglong, C.c_int, _dp, _dp]
    n = L.oracle_film_stock_knots(STOCKS[stock], spectral_bins, stock_seed, emulsion, logs.ctypes.data_as(_dp),
                                  fracs.ctypes.data_as(_dp))
    if n <= 0:
        raise ValueError("oracle film_stock_knots failed")
    return logs[:n], fracs[:n]


def film_stock_sensitivity(stock: str, emulsion: int, spectral_bins: int = 16, stock_seed: int = 0) -> np.ndarray:
    out = np.zeros(81)
    L = lib()
    L.oracle_film_stock_sensitivity.restype = C.c_int
    L.oracle_film_stock_sensitivity.argtypes = [C.c_int, C.c_int, C.c_ulonglong, C.c_int, _dp]
    n = L.oracle_film_stock_sensitivity(STOCKS[stock], spectral_bins, stock_seed, emulsion, out.ctypes.data_as(_dp))
    if n <= 0:
        raise ValueError("oracle film_stock_sensitivity failed")
    return out[:n]


def film_cmf(spectral_bins: int) -> np.ndarray:
    """CIE 1931 CMFs (3 x bins) on the grid: the reference's table at 16 bins, the fit at 81."""
    out = np.zeros(3 * spectral_bins)
    L = lib()
    L.oracle_film_cmf.restype = None
    L.oracle_film_cmf.argtypes = [C.c_int, _dp]
    L.oracle_film_cmf(spectral_bins, out.ctypes.data_as(_dp))
    return out.reshape(3, spectral_bins)


def film_capture_lut(stock: str, emulsion: int):
    """(log10 phi knots, developable fractions, (mu_ln, sigma_ln, k)) of one emulsion (capture.rs:37-106)."""
    logs, fracs, par = np.zeros(64), np.zeros(64), np.zeros(3)
    _rc(_film_sigs(lib()).oracle_film_capture_lut(STOCKS[stock], emulsion, logs.ctypes.data_as(_dp),
                                                   fracs.ctypes.data_as(_dp), par.ctypes.data_as(_dp)), "lut")
    return logs, fracs, tuple(par)


def film_lut_sample(stock: str, emulsion: int, phi: float) -> float:
    return _film_sigs(lib()).oracle_film_lut_sample(STOCKS[stock], emulsion, phi)


def film_upsample(rgb) -> np.ndarray:
    """upsample_acescg (upsample.rs:145-170): 16 spectral samples 400..700 nm."""
    a = np.asarray(rgb, np.float32)
    out = np.zeros(16)
    _film_sigs(lib()).oracle_film_upsample(_f(a), out.ctypes.data_as(_dp))
    return out


def film_spectrum_to_acescg(s) -> np.ndarray:
    """spectrum_to_acescg_rgb (upsample.rs:179-184)."""
    a = np.ascontiguousarray(s, np.float64)
    out = np.zeros(3)
    _film_sigs(lib()).oracle_film_spectrum_to_acescg(a.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    return out


def film_normalized_dmin(stock: str = "ColorNeg200") -> np.ndarray:
    """normalized_dmin_acescg (densitometry.rs:76-117)."""
    out = np.zeros(3, np.float32)
    _rc(_film_sigs(lib()).oracle_film_normalized_dmin(STOCKS[stock], _f(out)), "dmin")
    return out


def _fit_sigs(L):
    if not getattr(L, "_fit_sigs", False):
        L.oracle_film_fit_nparams.restype = C.c_int
        L.oracle_film_fit_nominal.restype = C.c_int
        L.oracle_film_fit_nominal.argtypes = [C.c_int, C.c_int, C.c_uint64, _dp]
        L.oracle_film_fit.restype = C.c_int
        L.oracle_film_fit.argtypes = [_fp, _sz, _sz, C.c_int, C.c_int, C.c_uint64, C.c_float, _dp, C.c_int]
        L.oracle_film_fit_gradient.restype = C.c_int
        L.oracle_film_fit_gradient.argtypes = [_fp, _fp, _sz, _sz, C.c_int, C.c_int, C.c_uint64, C.c_float, _dp, _dp,
                                               _fp]
        L._fit_sigs = True
    return L


# flat layout of the C API: n_emulsions, n_bins, n_knots, sens 8x81, knots 8x8, d_max 8, gamma 8,
# hal_strength, hal_radius
_FIT_SENS, _FIT_KNOTS = 3, 3 + 8 * 81
_FIT_DMAX = _FIT_KNOTS + 64
_FIT_GAMMA = _FIT_DMAX + 8
_FIT_HAL = _FIT_GAMMA + 8
FIT_NPARAMS = _FIT_HAL + 2


def _fit_to_flat(params: dict, E: int, nb: int, K: int) -> np.ndarray:
    v = np.zeros(FIT_NPARAMS)
    v[0:3] = (E, nb, K)
    s = np.zeros((8, 81))
    s[:E, :nb] = np.asarray(params["sens"], np.float64)
    v[_FIT_SENS:_FIT_KNOTS] = s.ravel()
    if K:
        k = np.zeros((8, 8))
        k[:E, :K] = np.asarray(params["knots"], np.float64)
        v[_FIT_KNOTS:_FIT_DMAX] = k.ravel()
    v[_FIT_DMAX:_FIT_DMAX + E] = np.asarray(params["d_max"], np.float64)
    v[_FIT_GAMMA:_FIT_GAMMA + 8] = 1.0
    v[_FIT_GAMMA:_FIT_GAMMA + E] = np.asarray(params["gamma"], np.float64)
    v[_FIT_HAL] = float(params["hal_strength"])
    v[_FIT_HAL + 1] = float(params["hal_radius"])
    return v


def _fit_from_flat(v: np.ndarray) -> dict:
    E, nb, K = int(v[0]), int(v[1]), int(v[2])
    d = {"sens": v[_FIT_SENS:_FIT_KNOTS].reshape(8, 81)[:E, :nb].copy(),
         "d_max": v[_FIT_DMAX:_FIT_DMAX + E].copy(), "gamma": v[_FIT_GAMMA:_FIT_GAMMA + E].copy(),
         "hal_strength": float(v[_FIT_HAL]), "hal_radius": float(v[_FIT_HAL + 1])}
    if K:
        d["knots"] = v[_FIT_KNOTS:_FIT_DMAX].reshape(8, 8)[:E, :K].copy()
    return d


def film_fit_nominal(stock: str = "ColorNeg200", bins: int = 16, stock_seed: int = 0) -> dict:
    """Nominal parameters of the differentiable film model: sens (E x bins), d_max, gamma,
    hal_strength, hal_radius, and knots (E x n_knots) for a knot-curve stock."""
    L = _fit_sigs(lib())
    assert L.oracle_film_fit_nparams() == FIT_NPARAMS
    v = np.zeros(FIT_NPARAMS)
    _rc(L.oracle_film_fit_nominal(STOCKS[stock], bins, stock_seed, v.ctypes.data_as(_dp)), "film_fit_nominal")
    return _fit_from_flat(v)


def _fit_flat_for(params: dict, stock: str, bins: int, stock_seed: int) -> np.ndarray:
    nom = film_fit_nominal(stock, bins, stock_seed)
    E = len(nom["d_max"])
    K = nom["knots"].shape[1] if "knots" in nom else 0
    return _fit_to_flat(params, E, bins, K)


def film_fit(rgb: np.ndarray, params: dict, stock: str = "ColorNeg200", pitch_um: float = 36000.0 / 6016,
             bins: int = 16, stock_seed: int = 0, fp64: bool = True) -> np.ndarray:
    """The differentiable film model (film::process with grain, DIR and adjacency off,
    parameters overridden, scan normalised by the nominal Dmin) on an (H, W, 3) image."""
    L = _fit_sigs(lib())
    out = np.array(rgb, np.float32, copy=True, order="C")
    h, w, _ = out.shape
    v = _fit_flat_for(params, stock, bins, stock_seed)
    _rc(L.oracle_film_fit(_f(out), w, h, STOCKS[stock], bins, stock_seed, pitch_um, v.ctypes.data_as(_dp),
                          int(fp64)), "film_fit")
    return out


def film_fit_gradient(rgb: np.ndarray, gout: np.ndarray, params: dict, stock: str = "ColorNeg200",
                      pitch_um: float = 36000.0 / 6016, bins: int = 16, stock_seed: int = 0):
    """FP64 reverse pass of the model: (dL/dparams as a dict, the f64 forward's output)
    for dL/dout = gout on one (H, W, 3) image."""
    L = _fit_sigs(lib())
    x = np.ascontiguousarray(rgb, np.float32)
    gy = np.ascontiguousarray(gout, np.float32)
    h, w, _ = x.shape
    v = _fit_flat_for(params, stock, bins, stock_seed)
    g = np.zeros(FIT_NPARAMS)
    out = np.zeros_like(x)
    _rc(L.oracle_film_fit_gradient(_f(x), _f(gy), w, h, STOCKS[stock], bins, stock_seed, pitch_um,
                                   v.ctypes.data_as(_dp), g.ctypes.data_as(_dp), _f(out)), "film_fit_gradient")
    return _fit_from_flat(g), out
