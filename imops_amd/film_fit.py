"""Differentiable film model on the GPU, for fitting film-stock parameters.

The model is the film stage (pichromatic/src/film/mod.rs:77-129) in its
NegativeLinear form with grain, DIR inhibition and adjacency switched off,
parameterised by the emulsion sensitivities (E x 16 samples, 400..700 nm), the
characteristic curve (d_max, gamma per emulsion) and two halation multipliers
(gain and PSF width).  Forward and reverse passes are the library's HIP
kernels (`imops_film_fit_forward` / `imops_film_fit_backward`); `FilmFit`
wraps them as a torch.autograd.Function so an optimiser can drive them.

    model = FilmModel(im.Context(0))
    params = model.nominal()                       # dict of float32 CPU tensors
    out = model(rgb, **params)                     # rgb: (N, H, W, 3) float32 CUDA, ACEScg absolute
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np
import torch

from . import Context, STOCKS
from ._lib import FilmFitParams, check, lib

PARAM_NAMES = ("sens", "d_max", "gamma", "hal_strength", "hal_radius")


class FilmModel:
    def __init__(self, ctx: Context, stock: str = "ColorNeg200", pitch_um: float = 36000.0 / 6016):
        self.ctx = ctx
        self.stock = STOCKS[stock]
        self.pitch = float(pitch_um)
        p = FilmFitParams()
        check(lib().imops_film_fit_nominal(self.stock, C.byref(p)))
        self.n_emulsions = p.n_emulsions
        self._lib_stream = None

    def _after_torch(self, device) -> None:
        """Order the library's stream after the work queued on torch's current stream
        (an event wait on the device, no host synchronisation)."""
        if self._lib_stream is None:
            self._lib_stream = torch.cuda.ExternalStream(self.ctx.stream, device=device)
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream(device))
        self._lib_stream.wait_event(ev)

    def _before_torch(self, device) -> None:
        """Order torch's current stream after the library's queued work."""
        ev = torch.cuda.Event()
        ev.record(self._lib_stream)
        torch.cuda.current_stream(device).wait_event(ev)

    def nominal(self) -> dict:
        p = FilmFitParams()
        check(lib().imops_film_fit_nominal(self.stock, C.byref(p)))
        E = p.n_emulsions
        return {"sens": torch.tensor(np.ctypeslib.as_array(p.sens)[:E].copy()),
                "d_max": torch.tensor(np.ctypeslib.as_array(p.d_max)[:E].copy()),
                "gamma": torch.tensor(np.ctypeslib.as_array(p.gamma)[:E].copy()),
                "hal_strength": torch.tensor(p.hal_strength, dtype=torch.float32),
                "hal_radius": torch.tensor(p.hal_radius, dtype=torch.float32)}

    def pack(self, sens, d_max, gamma, hal_strength, hal_radius) -> FilmFitParams:
        p = FilmFitParams()
        E = self.n_emulsions
        p.n_emulsions = E
        s = np.zeros((8, 16), np.float32)
        s[:E] = sens.detach().cpu().numpy()
        dm = np.zeros(8, np.float32)
        dm[:E] = d_max.detach().cpu().numpy()
        gm = np.ones(8, np.float32)
        gm[:E] = gamma.detach().cpu().numpy()
        C.memmove(p.sens, s.ctypes.data, s.nbytes)
        C.memmove(p.d_max, dm.ctypes.data, dm.nbytes)
        C.memmove(p.gamma, gm.ctypes.data, gm.nbytes)
        p.hal_strength = float(hal_strength)
        p.hal_radius = float(hal_radius)
        return p

    def state_like(self, rgb: torch.Tensor) -> torch.Tensor:
        """Device buffer for the forward state that a backward pass reuses."""
        n, h, w, _ = rgb.shape
        floats = C.c_size_t()
        check(lib().imops_film_fit_state_floats(self.stock, n, w, h, C.byref(floats)))
        return torch.empty(floats.value, dtype=torch.float32, device=rgb.device)

    def forward_raw(self, rgb: torch.Tensor, p: FilmFitParams, state: torch.Tensor | None = None) -> torch.Tensor:
        _check_images(rgb)
        n, h, w, _ = rgb.shape
        out = torch.empty_like(rgb)
        self._after_torch(rgb.device)
        check(lib().imops_film_fit_forward(self.ctx.handle, self.stock, self.pitch, C.byref(p), rgb.data_ptr(),
                                           out.data_ptr(), n, w, h, None if state is None else state.data_ptr()))
        self._before_torch(rgb.device)
        return out

    def backward_raw(self, rgb: torch.Tensor, grad_out: torch.Tensor, p: FilmFitParams,
                     state: torch.Tensor | None = None) -> FilmFitParams:
        """dL/dparams; `state` (from forward_raw on the same images and parameters) skips the
        recomputation of the exposure and blurs."""
        _check_images(rgb)
        grad_out = grad_out.contiguous()
        n, h, w, _ = rgb.shape
        g = FilmFitParams()
        self._after_torch(rgb.device)
        check(lib().imops_film_fit_backward(self.ctx.handle, self.stock, self.pitch, C.byref(p), rgb.data_ptr(),
                                            grad_out.data_ptr(), n, w, h, None if state is None else state.data_ptr(),
                                            C.byref(g)))  # returns after the device work (it reads the sums back)
        return g

    def __call__(self, rgb, sens, d_max, gamma, hal_strength, hal_radius) -> torch.Tensor:
        return FilmFit.apply(rgb, sens, d_max, gamma, hal_strength, hal_radius, self)


def _check_images(rgb: torch.Tensor) -> None:
    if not (rgb.is_cuda and rgb.dtype == torch.float32 and rgb.dim() == 4 and rgb.shape[-1] == 3
            and rgb.is_contiguous()):
        raise ValueError("film fit images must be a contiguous float32 CUDA tensor of shape (N, H, W, 3)")


class FilmFit(torch.autograd.Function):
    """out = film model(rgb; sens, d_max, gamma, hal_strength, hal_radius); gradients w.r.t. the parameters."""

    @staticmethod
    def forward(fctx, rgb, sens, d_max, gamma, hal_strength, hal_radius, model):
        p = model.pack(sens, d_max, gamma, hal_strength, hal_radius)
        fctx.model, fctx.packed = model, p
        # keep the forward's exposure planes for the backward when gradients are wanted
        state = model.state_like(rgb) if any(fctx.needs_input_grad[1:6]) else None
        out = model.forward_raw(rgb, p, state)
        fctx.save_for_backward(rgb, sens, d_max, gamma, hal_strength, hal_radius, state)
        return out

    @staticmethod
    def backward(fctx, grad_out):
        rgb, sens, d_max, gamma, hs, hr, state = fctx.saved_tensors
        g = fctx.model.backward_raw(rgb, grad_out, fctx.packed, state)
        E = fctx.model.n_emulsions
        as_t = lambda a, like: torch.tensor(np.asarray(a, np.float32), dtype=like.dtype, device=like.device)
        return (None, as_t(np.ctypeslib.as_array(g.sens)[:E], sens), as_t(np.ctypeslib.as_array(g.d_max)[:E], d_max),
                as_t(np.ctypeslib.as_array(g.gamma)[:E], gamma), as_t(g.hal_strength, hs).reshape(hs.shape),
                as_t(g.hal_radius, hr).reshape(hr.shape), None)


def fit(model: FilmModel, rgb: torch.Tensor, target: torch.Tensor, init: dict, steps: int = 200, lr: float = 0.01,
        log_every: int = 0) -> dict:
    """Adam on the L2 loss between the model output and `target`; returns the fitted
    parameters, the loss history and the mean forward+backward time per step (ms)."""
    params = {k: v.detach().clone().requires_grad_(True) for k, v in init.items()}
    opt = torch.optim.Adam(params.values(), lr=lr)
    history, times = [], []
    for it in range(steps):
        t0 = time.perf_counter()
        opt.zero_grad()
        out = model(rgb, **params)
        loss = ((out - target) ** 2).sum()
        loss.backward()
        times.append(time.perf_counter() - t0)
        opt.step()
        with torch.no_grad():
            params["sens"].clamp_(min=0.0)
            params["gamma"].clamp_(min=0.05)
            params["d_max"].clamp_(min=0.05)
            params["hal_strength"].clamp_(min=0.0)
            params["hal_radius"].clamp_(min=0.05)
        history.append(loss.detach())  # read back once at the end, not every step
        if log_every and it % log_every == 0:
            print(f"[fit] step {it} loss {float(history[-1]):.6g}", flush=True)
    history = torch.stack(history).cpu().tolist() if history else []
    return {"params": {k: v.detach() for k, v in params.items()}, "loss": history,
            "ms_per_iter": 1e3 * float(np.median(times))}
