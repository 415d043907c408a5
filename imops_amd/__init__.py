"""MI355X (gfx950) implementation of the pichromatic raw -> demosaic -> BM3D ->
colour -> film stage chain.

The Python layer mirrors the reference's host interface for this path:

* :class:`Image` follows ``pichromatic::pixel::Image`` and its stage methods
  (pichromatic/src/pixel.rs:55-171), but its pixel buffers live in device
  memory behind a handle of the C library;
* :func:`parse_config` / :func:`run_pixel_pipeline` follow
  ``config::parse_config`` (pichromatic-pipeline/src/config.rs:13-34) and
  ``run_pixel_pipeline`` (pichromatic-pipeline/src/pipeline.rs:4-13).

Every stage runs as hand-written HIP kernels in ``libimops_gpu.so``; a missing
library raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import (CS_ACESCG, CS_LINEAR_SRGB, CS_NONE, CS_SRGB, CS_XYZD65, DEMOSAIC_AMAZE, DEMOSAIC_MARKESTEIJN,
                   ImopsError, check, fptr, lib)

__all__ = [
    "Context", "Image", "PipelineConfig", "parse_config", "run_pixel_pipeline", "ImopsError", "synth_raw_u16",
    "synth_scene_rgb", "CS_NONE", "CS_SRGB", "CS_LINEAR_SRGB", "CS_ACESCG", "CS_XYZD65", "DEMOSAIC_AMAZE",
    "DEMOSAIC_MARKESTEIJN", "COLOR_SPACES", "bm3d_params",
]

# ColorSpaceTag names accepted by CST (modules.rs:107-112, serde_plain names)
COLOR_SPACES = {"Srgb": CS_SRGB, "LinearSrgb": CS_LINEAR_SRGB, "AcesCg": CS_ACESCG, "XyzD65": CS_XYZD65}
_ALGORITHMS = {"amaze": DEMOSAIC_AMAZE, "markesteijn": DEMOSAIC_MARKESTEIJN, "fast": _lib.DEMOSAIC_FAST,
               "superfast": _lib.DEMOSAIC_SUPERFAST, "supersuperfast": _lib.DEMOSAIC_SUPERSUPERFAST,
               "markesteijn3": _lib.DEMOSAIC_MARKESTEIJN3, "markesteijn3_1pass": _lib.DEMOSAIC_MARKESTEIJN3_1PASS}
# Film stage choices (modules.rs:800-830; StockId order of film/stock/mod.rs:124-133)
STOCKS = {"BwStub": 0, "ColorNeg200": 1, "Portra400": 2, "Ektar100": 3, "FujiPro400H": 4, "EktachromeE100": 5,
          "TriX400": 6, "Synthetic": 7}
FILM_FORMATS = {"Film35mm": 0, "Film6x6": 1, "Film4x5": 2}
FILM_OUTPUTS = {"NegativeLinear": 0, "PositiveLinear": 1}


def _cs(target) -> int:
    if isinstance(target, str):
        if target not in COLOR_SPACES:
            raise ImopsError(_lib.E_INVALID, f"Color space not recognized: {target}")
        return COLOR_SPACES[target]
    return int(target)


class Context:
    """One device + one HIP stream (imops_ctx)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().imops_ctx_create(device, C.byref(h)))
        self._h = h
        self.device = device

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    @property
    def stream(self) -> int:
        return lib().imops_ctx_stream(self._h) or 0

    def synchronize(self) -> None:
        check(lib().imops_ctx_synchronize(self._h))

    def set_timing(self, enabled: bool) -> None:
        """Record HIP events around the library's launches (clears earlier records)."""
        check(lib().imops_ctx_set_timing(self._h, int(bool(enabled))))

    def timing(self, region: str):
        """(total ms, launches) recorded for a region since timing was enabled."""
        ms, n = C.c_double(), C.c_int()
        check(lib().imops_ctx_timing(self._h, region.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def set_autotune(self, enabled: bool) -> None:
        """Time every launch-shape candidate of a tunable kernel on its first call at a
        new shape and keep the fastest (results are identical for every candidate)."""
        check(lib().imops_ctx_set_autotune(self._h, int(bool(enabled))))

    def measure_peaks(self) -> dict:
        """Measured rates of this context's device (HBM copy, FP32 non-fused VALU issue,
        FP16 MFMA); allocates two 1 GiB buffers for the call."""
        out = (C.c_double * 6)()
        check(lib().imops_measure_peaks(self._h, out, 6))
        return {"hbm_copy_GBps": round(out[0], 1), "valu_fp32_nonfused_instr_per_clk_per_simd": round(out[1], 4),
                "valu_fp32_nonfused_TFLOPs": round(out[2], 2), "mfma_f16_TFLOPs": round(out[3], 1),
                "rated_clock_GHz": round(out[4], 3), "compute_units": int(out[5])}

    def tuning(self) -> dict:
        """Launch-shape choices made so far: {key: value}."""
        need = C.c_size_t()
        check(lib().imops_ctx_tuning(self._h, None, 0, C.byref(need)))
        buf = C.create_string_buffer(need.value)
        check(lib().imops_ctx_tuning(self._h, buf, need.value, None))
        out = {}
        for line in buf.value.decode().splitlines():
            k, _, v = line.partition("=")
            out[k] = int(v)
        return out

    def close(self) -> None:
        if self._h:
            lib().imops_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Image:
    """Device-resident image (raw CFA plane and/or interleaved RGB f32 + metadata)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        h = C.c_void_p()
        check(lib().imops_image_create(ctx.handle, C.byref(h)))
        self._h = h

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def close(self) -> None:
        if self._h:
            lib().imops_image_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- ingest / readback -------------------------------------------------
    def set_raw_u16(self, raw, cfa: str, black: float, white: float, crop=None) -> "Image":
        """u16 sensor frame (numpy host array, or a device tensor) -> normalised CFA plane."""
        on_dev, ptr, (h, w) = _buffer(raw, np.uint16)
        cx, cy, cw, ch = crop if crop is not None else (0, 0, w, h)
        check(lib().imops_image_set_raw_u16(self.ctx.handle, self._h, ptr, on_dev, w, h, cx, cy, cw, ch,
                                            float(black), float(white), cfa.encode()))
        return self

    def bind_raw_u16(self, raw, cfa: str, black: float, white: float, crop=None) -> "Image":
        """Binds a device u16 frame (a torch tensor kept alive by the caller until the
        demosaic has run) without copying: AMaZE normalises it in its tile fill."""
        on_dev, ptr, (h, w) = _buffer(raw, np.uint16)
        if not on_dev:
            raise ValueError("bind_raw_u16: the frame must be a device tensor")
        cx, cy, cw, ch = crop if crop is not None else (0, 0, w, h)
        self._bound = raw  # keeps the tensor alive while bound
        check(lib().imops_image_bind_raw_u16(self.ctx.handle, self._h, ptr, w, h, cx, cy, cw, ch, float(black),
                                             float(white), cfa.encode()))
        return self

    def crop_bayer_center(self, crop_factor: int) -> "Image":
        """crop_bayer_center (extern_pipeline.rs:113-179) on the raw plane."""
        check(lib().imops_crop_bayer_center(self.ctx.handle, self._h, int(crop_factor)))
        return self

    def set_raw_f32(self, raw, cfa: str, crop_origin=(0, 0)) -> "Image":
        on_dev, ptr, (h, w) = _buffer(raw, np.float32)
        check(lib().imops_image_set_raw_f32(self.ctx.handle, self._h, ptr, on_dev, w, h, cfa.encode(),
                                            crop_origin[0], crop_origin[1]))
        return self

    def set_rgb(self, rgb, color_space=CS_NONE) -> "Image":
        on_dev, ptr, shape = _buffer(rgb, np.float32)
        if len(shape) != 3 or shape[2] != 3:
            raise ImopsError(_lib.E_INVALID, "rgb must have shape (H, W, 3)")
        check(lib().imops_image_set_rgb_f32(self.ctx.handle, self._h, ptr, on_dev, shape[1], shape[0],
                                            _cs(color_space)))
        return self

    def info(self):
        w, h, cs, has = C.c_size_t(), C.c_size_t(), C.c_int(), C.c_int()
        check(lib().imops_image_get_info(self._h, C.byref(w), C.byref(h), C.byref(cs), C.byref(has)))
        return w.value, h.value, cs.value, bool(has.value)

    @property
    def width(self) -> int:
        return self.info()[0]

    @property
    def height(self) -> int:
        return self.info()[1]

    @property
    def color_space(self) -> int:
        return self.info()[2]

    def rgb(self) -> np.ndarray:
        w, h, _, _ = self.info()
        out = np.empty((h, w, 3), np.float32)
        check(lib().imops_image_get_rgb_f32(self.ctx.handle, self._h, out.ctypes.data_as(C.c_void_p), 0, out.size))
        return out

    def _encoded(self, fn, dtype, orientation: int) -> np.ndarray:
        w, h, _, _ = self.info()
        shape = (w, h, 3) if orientation in (1, 3) else (h, w, 3)
        out = np.empty(shape, dtype)
        check(fn(self.ctx.handle, self._h, out.ctypes.data_as(C.c_void_p), 0, int(orientation), out.size))
        return out

    def rgb_u8(self, orientation: int = 0) -> np.ndarray:
        """8-bit RGB as get_image_rgb_data_c encodes it (extern_pipeline.rs:362-395), oriented
        (0 none, 1 Rotate90, 2 Rotate180, 3 Rotate270; output.rs:31-70)."""
        return self._encoded(lib().imops_image_get_rgb_u8, np.uint8, orientation)

    def rgb_f16(self, orientation: int = 0) -> np.ndarray:
        """Half-float RGB as save_exr writes it (output.rs:201-209), oriented as rgb_u8."""
        return self._encoded(lib().imops_image_get_rgb_f16, np.uint16, orientation).view(np.float16)

    def rgb_device_ptr(self) -> int:
        p = C.c_void_p()
        check(lib().imops_image_rgb_device_ptr(self._h, C.byref(p)))
        return p.value or 0

    # -- metadata ----------------------------------------------------------
    def set_wb_coeffs(self, wb: Sequence[float]) -> "Image":
        a = np.asarray(wb, np.float32)
        if a.size != 4:
            raise ImopsError(_lib.E_INVALID, "wb_coeffs has 4 entries")
        check(lib().imops_image_set_wb_coeffs(self._h, fptr(a)))
        return self

    def set_calibration_matrix(self, m: Sequence[float]) -> "Image":
        a = np.ascontiguousarray(m, np.float32).ravel()
        check(lib().imops_image_set_calibration_matrix(self._h, fptr(a), a.size))
        return self

    def set_shutter_seconds(self, s: float) -> "Image":
        check(lib().imops_image_set_shutter_seconds(self._h, float(s)))
        return self

    def set_exposure_metadata(self, baseline_ev: float, f_number: float, iso: float) -> "Image":
        check(lib().imops_image_set_exposure_metadata(self._h, float(baseline_ev), float(f_number), float(iso)))
        return self

    def set_opcode_list3(self, data: bytes) -> "Image":
        buf = C.create_string_buffer(bytes(data), len(data))
        check(lib().imops_image_set_opcode_list3(self._h, C.cast(buf, C.c_void_p), len(data)))
        return self

    # -- stages (pixel.rs:55-171) -------------------------------------------
    def demosaic(self, algorithm="amaze") -> "Image":
        a = _ALGORITHMS[algorithm] if isinstance(algorithm, str) else int(algorithm)
        check(lib().imops_demosaic(self.ctx.handle, self._h, a))
        return self

    def cfa_coeffs(self, wb: Optional[Sequence[float]] = None) -> "Image":
        if wb is None:
            check(lib().imops_cfa_coeffs(self.ctx.handle, self._h, None))
        else:
            a = np.asarray(wb, np.float32)
            check(lib().imops_cfa_coeffs(self.ctx.handle, self._h, fptr(a)))
        return self

    def exp(self, ev: float) -> "Image":
        check(lib().imops_exp(self.ctx.handle, self._h, float(ev)))
        return self

    def gamma(self, gamma: float) -> "Image":
        check(lib().imops_gamma(self.ctx.handle, self._h, float(gamma)))
        return self

    def cst(self, target) -> "Image":
        check(lib().imops_cst(self.ctx.handle, self._h, _cs(target)))
        return self

    def bm3d(self, intensity: float = 0.0, **explicit) -> "Image":
        """BM3D (bm3d.rs:34-61).  Keyword overrides: sigma (0..255 units), lambda_, max_dist,
        chroma_scale, detail, stride, search_window, max_group, max_candidates, and
        group_transform ("hadamard", the reference's; "haar"), and
        precision ("fp32"; "mixed-match": FP16 matrix-core matching, same groups, FP32 filter;
        "mixed": that matcher plus the filter's transforms on FP16 matrix cores)."""
        if not explicit:
            check(lib().imops_bm3d(self.ctx.handle, self._h, float(intensity)))
            return self
        p = bm3d_params(intensity, **explicit)
        check(lib().imops_bm3d_ex(self.ctx.handle, self._h, C.byref(p)))
        return self

    def chroma_bm3d(self, intensity: float = 0.0) -> "Image":
        """Chroma-only denoise (Image::chroma_bm3d, pixel.rs:122-126; bm3d.rs:705-759)."""
        check(lib().imops_chroma_bm3d(self.ctx.handle, self._h, float(intensity)))
        return self

    def chroma_denoise(self, radius: int = 4, epsilon: float = 0.01) -> "Image":
        """Luma-guided chroma denoise (Image::chroma_denoise, pixel.rs:128-137; chroma_denoise.rs:14-152)."""
        check(lib().imops_luma_chroma_denoise(self.ctx.handle, self._h, int(radius), float(epsilon)))
        return self

    def contrast(self, c: float) -> "Image":
        """Image::contrast (pixel.rs:107-110)."""
        check(lib().imops_contrast(self.ctx.handle, self._h, float(c)))
        return self

    def highlight_reconstruction(self) -> "Image":
        """Image::highlight_reconstruction (pixel.rs:111-114) with the image's white balance."""
        check(lib().imops_highlight_reconstruction(self.ctx.handle, self._h))
        return self

    def lch(self, lc: float = 1.0, cc: float = 1.0, hc: float = 1.0) -> "Image":
        """Image::lch (pixel.rs:60-68)."""
        check(lib().imops_lch(self.ctx.handle, self._h, float(lc), float(cc), float(hc)))
        return self

    def sigmoid_tone_map(self) -> "Image":
        """Image::sigmoid_tone_map (pixel.rs:91-94)."""
        check(lib().imops_sigmoid_tone_map(self.ctx.handle, self._h))
        return self

    def baseline_exposure(self) -> "Image":
        """BaselineExposureCompensation (modules.rs:332-371) from the image's exposure metadata."""
        check(lib().imops_baseline_exposure(self.ctx.handle, self._h))
        return self

    def vignette(self, strength: float = 1.0) -> "Image":
        """Image::vignette (pixel.rs:149-160) from the image's OpcodeList3."""
        check(lib().imops_vignette(self.ctx.handle, self._h, float(strength)))
        return self

    def film(self, stock: str = "ColorNeg200", film_format: str = "Film35mm", seed: int = 1,
             output: str = "NegativeLinear", **ext) -> "Image":
        """Film simulation (film/mod.rs:77-129); the image must be absolute-luminance ACEScg.

        Extension keywords (imops_film_ex; absent: the reference's behaviour): precision
        ("fp64" | "mixed"), spectral_bins (16 | 81), stock_seed (stock "Synthetic"),
        grain ("splitmix" | "philox"), grain_sigma, halation_radius, halation_strength."""
        if not ext:
            check(lib().imops_film(self.ctx.handle, self._h, STOCKS[stock], FILM_FORMATS[film_format], int(seed),
                                   FILM_OUTPUTS[output]))
            return self
        p = _lib.FilmParams()
        check(lib().imops_film_default_params(C.byref(p)))
        p.stock, p.film_format, p.seed, p.output = STOCKS[stock], FILM_FORMATS[film_format], int(seed), FILM_OUTPUTS[output]
        for k, v in ext.items():
            if k == "precision":
                p.precision = {"fp64": 0, "mixed": 1}[v]
            elif k == "grain":
                p.grain = {"splitmix": 0, "philox": 1}[v]
            elif k in ("spectral_bins", "stock_seed"):
                setattr(p, k, int(v))
            elif k in ("grain_sigma", "halation_radius", "halation_strength"):
                setattr(p, k, float(v))
            else:
                raise TypeError(f"film: unknown parameter {k}")
        check(lib().imops_film_ex(self.ctx.handle, self._h, C.byref(p)))
        return self

    def camera_cst(self, target, calibration_matrix: Sequence[float]) -> "Image":
        a = np.ascontiguousarray(calibration_matrix, np.float32).ravel()
        check(lib().imops_camera_cst(self.ctx.handle, self._h, _cs(target), fptr(a), a.size))
        return self


class PipelineConfig:
    """Parsed stage list (imops_pipeline)."""

    def __init__(self, text: str):
        h = C.c_void_p()
        data = text.encode()
        check(lib().imops_pipeline_parse(data, len(data), C.byref(h)))
        self._h = h

    @property
    def pipeline_modules(self) -> list[str]:
        n = lib().imops_pipeline_num_stages(self._h)
        return [lib().imops_pipeline_stage_name(self._h, i).decode() for i in range(n)]

    def __del__(self):
        try:
            if self._h:
                lib().imops_pipeline_free(self._h)
        except Exception:
            pass


def bm3d_params(intensity: float = 0.0, **explicit) -> _lib.Bm3dParams:
    """Bm3dParams::from_intensity (bm3d.rs:22-32) with optional explicit overrides."""
    p = _lib.Bm3dParams()
    check(lib().imops_bm3d_default_params(float(intensity), C.byref(p)))
    for k, v in explicit.items():
        if k == "precision":
            v = {"fp32": _lib.BM3D_FP32, "mixed": _lib.BM3D_MIXED,
                     "mixed-match": _lib.BM3D_MIXED_MATCH}[v] if isinstance(v, str) else int(v)
        elif k == "group_transform":
            v = {"hadamard": _lib.BM3D_HADAMARD, "haar": _lib.BM3D_HAAR}[v] if isinstance(v, str) else int(v)
        setattr(p, "lambda_" if k == "lambda" else k, v)
    return p


def parse_config(text: str) -> PipelineConfig:
    return PipelineConfig(text)


def run_pixel_pipeline(image: Image, config: PipelineConfig) -> Image:
    check(lib().imops_pipeline_run(image.ctx.handle, config._h, image.handle))
    return image


def dng_calibration_matrix(color_matrix_a=None, color_matrix_d65=None, color_matrix_other=None,
                           wb_coeffs=None) -> np.ndarray:
    """The calibration matrix parse_raw_image picks or interpolates (extern_pipeline.rs:254-270, 457-495)."""
    def arg(m, k):
        if m is None:
            return None, None
        a = np.ascontiguousarray(m, np.float32).ravel()
        if a.size != k:
            raise ValueError(f"expected {k} values")
        return a, fptr(a)
    keep = [arg(color_matrix_a, 9), arg(color_matrix_d65, 9), arg(color_matrix_other, 9), arg(wb_coeffs, 4)]
    out = np.zeros(9, np.float32)
    check(lib().imops_dng_calibration_matrix(keep[0][1], keep[1][1], keep[2][1], keep[3][1], fptr(out)))
    return out


def demosaic_amaze_noalloc(ctx: Context, cfa_plane_ptr: int, rgb_out_ptr: int, width: int, height: int,
                           cfa: str = "RGGB", crop_origin=(0, 0)) -> None:
    """AMaZE from a caller-owned device CFA plane into a caller-owned device RGB buffer
    (raw device addresses, e.g. torch's data_ptr()), allocating nothing; enqueued on the
    context's stream.  See imops_demosaic_amaze_noalloc in include/imops_gpu.h."""
    check(lib().imops_demosaic_amaze_noalloc(ctx.handle, cfa_plane_ptr, rgb_out_ptr, width, height, cfa.encode(),
                                             crop_origin[0], crop_origin[1]))


def run_burst(frames: list, config: PipelineConfig, graphs: bool = True) -> list:
    """The stage list over every frame (imops_pipeline_run_burst); frames of one
    shape after the first replay a HIP graph captured on a staging image."""
    if not frames:
        return frames
    ctx = frames[0].ctx
    arr = (C.c_void_p * len(frames))(*[f.handle for f in frames])
    check(lib().imops_pipeline_run_burst(ctx.handle, config._h, arr, len(frames), 1 if graphs else 0))
    return frames


def synth_scene_rgb(width: int, height: int, frame: int = 0) -> np.ndarray:
    out = np.empty((height, width, 3), np.float32)
    check(lib().imops_synth_scene_rgb(fptr(out), width, height, frame))
    return out


def synth_raw_u16(width: int, height: int, cfa: str = "RGGB", frame: int = 0, full_well: float = 20000.0,
                  black: float = 600.0, white: float = 16383.0, read_sigma: float = 2.0) -> np.ndarray:
    out = np.empty((height, width), np.uint16)
    check(lib().imops_synth_raw_u16(out.ctypes.data_as(C.c_void_p), width, height, cfa.encode(), frame,
                                    full_well, black, white, read_sigma))
    return out


def _buffer(x, dtype):
    """(on_device, pointer, shape) for a numpy array (host) or a torch CUDA tensor (device)."""
    if isinstance(x, np.ndarray):
        if x.dtype != dtype or not x.flags.c_contiguous:
            raise ImopsError(_lib.E_INVALID, f"expected a C-contiguous {np.dtype(dtype).name} array")
        return 0, C.c_void_p(x.ctypes.data), x.shape
    try:  # torch tensor
        import torch
    except ImportError as e:  # pragma: no cover
        raise ImopsError(_lib.E_INVALID, "unsupported buffer type") from e
    if isinstance(x, torch.Tensor):
        # u16 data may arrive as an int16 tensor (same bits; torch has few uint16 ops)
        want = {np.uint16: (torch.uint16, torch.int16), np.float32: (torch.float32,)}[dtype]
        if x.dtype not in want or not x.is_contiguous():
            raise ImopsError(_lib.E_INVALID, f"expected a contiguous {want[0]} tensor")
        return (1 if x.is_cuda else 0), C.c_void_p(x.data_ptr()), tuple(x.shape)
    raise ImopsError(_lib.E_INVALID, f"unsupported buffer type {type(x)!r}")
