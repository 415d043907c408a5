"""Differentiable film model on the GPU, for fitting film-stock parameters.

The model is the film stage (pichromatic/src/film/mod.rs:77-129) in its
NegativeLinear form with grain, DIR inhibition and adjacency switched off,
parameterised by the emulsion sensitivities (E x 16 samples, 400..700 nm), the
characteristic curve (d_max, gamma per emulsion) and two halation multipliers
(gain and PSF width).  Forward and reverse passes are the library's HIP
kernels (`imops_film_fit_forward` / `imops_film_fit_backward`); `FilmFit`
wraps them as a torch.autograd.Function so an optimiser can drive them.
`fit` runs Adam on the L2 loss through `imops_film_fit_l2_step` (loss and
gradient in one device pass per iteration) with the update on the host.

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

    def l2_step(self, rgb: torch.Tensor, target: torch.Tensor, p: FilmFitParams,
                grad: FilmFitParams | None = None, order: bool = True) -> tuple[float, FilmFitParams]:
        """L = sum (model(rgb) - target)^2 and dL/dparams in one device pass
        (imops_film_fit_l2_step: the model output is evaluated inside the reverse kernel,
        never written).  `grad`, when given, is filled in place.  order=False skips
        ordering the library's stream after torch's (the caller has done it once for
        inputs that no torch work changes since)."""
        _check_images(rgb)
        _check_images(target)
        if target.shape != rgb.shape:
            raise ValueError("film fit target must have the images' shape")
        n, h, w, _ = rgb.shape
        g = FilmFitParams() if grad is None else grad
        loss = C.c_double()
        if order or self._lib_stream is None:
            self._after_torch(rgb.device)
        check(lib().imops_film_fit_l2_step(self.ctx.handle, self.stock, self.pitch, C.byref(p), rgb.data_ptr(),
                                           target.data_ptr(), n, w, h, C.byref(g), C.byref(loss)))
        return loss.value, g

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


# lower bounds of the parameters after every optimiser step
LOWER = {"sens": 0.0, "d_max": 0.05, "gamma": 0.05, "hal_strength": 0.0, "hal_radius": 0.05}


class Adam:
    """torch.optim.Adam's update (default betas and eps, no weight decay) on one flat
    float32 numpy vector: m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2,
    theta -= lr / (1 - b1^t) * m / (sqrt(v) / sqrt(1 - b2^t) + eps)."""

    def __init__(self, theta: np.ndarray, lr: float, betas=(0.9, 0.999), eps: float = 1e-8):
        self.theta = theta
        self.lr, self.b1, self.b2, self.eps = np.float32(lr), betas[0], betas[1], np.float32(eps)
        self.m = np.zeros_like(theta)
        self.v = np.zeros_like(theta)
        self.t = 0

    def step(self, g: np.ndarray) -> None:
        self.t += 1
        self.m += (g - self.m) * np.float32(1.0 - self.b1)
        self.v *= np.float32(self.b2)
        self.v += g * g * np.float32(1.0 - self.b2)
        bc1 = 1.0 - self.b1 ** self.t
        bc2_sqrt = np.float32((1.0 - self.b2 ** self.t) ** 0.5)
        denom = np.sqrt(self.v) / bc2_sqrt + self.eps
        self.theta -= np.float32(self.lr / bc1) * self.m / denom


class FlatParams:
    """The fitted parameters as one float32 vector (sens E x 16, d_max E, gamma E,
    hal_strength, hal_radius) packed in place into one imops_film_fit_params, and the
    gradient struct read back as a vector of the same layout."""

    def __init__(self, model: FilmModel, init: dict):
        E = self.E = model.n_emulsions
        self.theta = np.concatenate([np.asarray(init[k].detach().cpu().numpy(), np.float32).ravel()
                                     for k in PARAM_NAMES])
        self.lower = np.concatenate([np.full(E * 16, LOWER["sens"]), np.full(E, LOWER["d_max"]),
                                     np.full(E, LOWER["gamma"]), [LOWER["hal_strength"], LOWER["hal_radius"]]]
                                    ).astype(np.float32)
        self.p = model.pack(**init)
        self.g = FilmFitParams()
        self._p_sens = np.ctypeslib.as_array(self.p.sens)[:E]
        self._p_dmax = np.ctypeslib.as_array(self.p.d_max)[:E]
        self._p_gamma = np.ctypeslib.as_array(self.p.gamma)[:E]
        self._g_sens = np.ctypeslib.as_array(self.g.sens)[:E]
        self._g_dmax = np.ctypeslib.as_array(self.g.d_max)[:E]
        self._g_gamma = np.ctypeslib.as_array(self.g.gamma)[:E]
        self.grad = np.zeros_like(self.theta)

    def pack(self) -> FilmFitParams:
        E, t = self.E, self.theta
        self._p_sens[:] = t[:16 * E].reshape(E, 16)
        self._p_dmax[:] = t[16 * E:17 * E]
        self._p_gamma[:] = t[17 * E:18 * E]
        self.p.hal_strength = float(t[18 * E])
        self.p.hal_radius = float(t[18 * E + 1])
        return self.p

    def read_grad(self) -> np.ndarray:
        E, g = self.E, self.grad
        g[:16 * E] = self._g_sens.ravel()
        g[16 * E:17 * E] = self._g_dmax
        g[17 * E:18 * E] = self._g_gamma
        g[18 * E] = self.g.hal_strength
        g[18 * E + 1] = self.g.hal_radius
        return g

    def as_dict(self) -> dict:
        E, t = self.E, self.theta
        return {"sens": torch.from_numpy(t[:16 * E].reshape(E, 16).copy()),
                "d_max": torch.from_numpy(t[16 * E:17 * E].copy()),
                "gamma": torch.from_numpy(t[17 * E:18 * E].copy()),
                "hal_strength": torch.tensor(float(t[18 * E])),
                "hal_radius": torch.tensor(float(t[18 * E + 1]))}


def fit(model: FilmModel, rgb: torch.Tensor, target: torch.Tensor, init: dict, steps: int = 200, lr: float = 0.01,
        log_every: int = 0, fused: bool = True) -> dict:
    """Adam on the L2 loss between the model output and `target`, the parameters clamped
    to LOWER after each step; returns the fitted parameters, the loss history (the loss
    at each step's parameters before its update) and the median wall time of one
    iteration (ms; forward, adjoint and optimiser step).

    fused: each iteration is one imops_film_fit_l2_step (loss and gradient in one device
    pass, one synchronisation) followed by Adam on the host on a 16E + 2E + 2 vector;
    otherwise the forward, a torch loss, autograd through FilmFit and torch.optim.Adam."""
    if not fused:
        return _fit_autograd(model, rgb, target, init, steps, lr, log_every)
    fp = FlatParams(model, init)
    opt = Adam(fp.theta, lr)
    history, times = [], []
    for it in range(steps):
        t0 = time.perf_counter()
        loss, _ = model.l2_step(rgb, target, fp.pack(), fp.g, order=it == 0)
        opt.step(fp.read_grad())
        np.maximum(fp.theta, fp.lower, out=fp.theta)
        times.append(time.perf_counter() - t0)
        history.append(loss)
        if log_every and it % log_every == 0:
            print(f"[fit] step {it} loss {loss:.6g}", flush=True)
    return {"params": fp.as_dict(), "loss": history, "ms_per_iter": 1e3 * float(np.median(times)) if times else 0.0}


def _fit_autograd(model, rgb, target, init, steps, lr, log_every) -> dict:
    params = {k: v.detach().clone().requires_grad_(True) for k, v in init.items()}
    opt = torch.optim.Adam(params.values(), lr=lr)
    history, times = [], []
    for it in range(steps):
        t0 = time.perf_counter()
        opt.zero_grad()
        out = model(rgb, **params)
        loss = ((out - target) ** 2).sum()
        loss.backward()
        opt.step()
        with torch.no_grad():
            for k, v in params.items():
                v.clamp_(min=LOWER[k])
        times.append(time.perf_counter() - t0)
        history.append(loss.detach())  # read back once at the end, not every step
        if log_every and it % log_every == 0:
            print(f"[fit] step {it} loss {float(history[-1]):.6g}", flush=True)
    history = torch.stack(history).cpu().tolist() if history else []
    return {"params": {k: v.detach() for k, v in params.items()}, "loss": history,
            "ms_per_iter": 1e3 * float(np.median(times)) if times else 0.0}
