"""Differentiable film model on the GPU, for fitting film-stock parameters.

The model is the film stage (pichromatic/src/film/mod.rs:77-129) in its
NegativeLinear form with grain, DIR inhibition and adjacency switched off, on
the stock's spectral grid (16 samples 400..700 nm as the reference, or 81
samples 380..780 nm), parameterised by the emulsion sensitivities on that grid,
the characteristic curve (d_max, gamma per emulsion, and for a knot-curve stock
such as the seeded synthetic one the developable fraction at each of its
knots) and two halation multipliers (gain and PSF width).  Forward and reverse
passes are the library's HIP kernels (`imops_film_fit_forward` /
`imops_film_fit_backward`); `FilmFit` wraps them as a torch.autograd.Function
so an optimiser can drive them.  `fit` runs Adam on the L2 loss through
`imops_film_fit_l2_step` (loss and gradient in one device pass per iteration)
with the update on the host.

    model = FilmModel(im.Context(0))                          # ColorNeg200, 16 bins
    model = FilmModel(ctx, "Synthetic", spectral_bins=81, stock_seed=7)  # 3 x 81 sensitivities, 3 x 5 knots
    params = model.nominal()                       # dict of float32 CPU tensors
    out = model(rgb, **params)                     # rgb: (N, H, W, 3) float32 CUDA, ACEScg absolute
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np
import torch

from . import Context, STOCKS
from ._lib import FilmFitModel, FilmFitParams, check, lib

PARAM_NAMES = ("sens", "d_max", "gamma", "hal_strength", "hal_radius", "knots")


class FilmModel:
    def __init__(self, ctx: Context, stock: str = "ColorNeg200", pitch_um: float = 36000.0 / 6016,
                 spectral_bins: int = 16, stock_seed: int = 0):
        self.ctx = ctx
        self.model = FilmFitModel(STOCKS[stock], int(spectral_bins), int(stock_seed), float(pitch_um))
        p = FilmFitParams()
        check(lib().imops_film_fit_nominal(C.byref(self.model), C.byref(p)))
        self.n_emulsions, self.n_bins, self.n_knots = p.n_emulsions, p.n_bins, p.n_knots
        self.names = PARAM_NAMES if self.n_knots else PARAM_NAMES[:5]
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
        check(lib().imops_film_fit_nominal(C.byref(self.model), C.byref(p)))
        return self.unpack(p)

    def knot_positions(self) -> np.ndarray:
        """log10 fluence of every knot (E x n_knots; empty for crystal-LUT stocks)."""
        out = np.zeros((self.n_emulsions, self.n_knots))
        for e in range(self.n_emulsions):
            row = np.zeros(max(self.n_knots, 1))
            check(lib().imops_film_fit_knots(C.byref(self.model), e, row.ctypes.data_as(C.POINTER(C.c_double))))
            out[e] = row[:self.n_knots]
        return out

    def unpack(self, p: FilmFitParams) -> dict:
        """A parameter (or gradient) struct as a dict of float32 CPU tensors."""
        E, nb, K = self.n_emulsions, self.n_bins, self.n_knots
        d = {"sens": torch.tensor(np.ctypeslib.as_array(p.sens)[:E, :nb].copy()),
             "d_max": torch.tensor(np.ctypeslib.as_array(p.d_max)[:E].copy()),
             "gamma": torch.tensor(np.ctypeslib.as_array(p.gamma)[:E].copy()),
             "hal_strength": torch.tensor(p.hal_strength, dtype=torch.float32),
             "hal_radius": torch.tensor(p.hal_radius, dtype=torch.float32)}
        if K:
            d["knots"] = torch.tensor(np.ctypeslib.as_array(p.knot_frac)[:E, :K].copy())
        return d

    def pack(self, sens, d_max, gamma, hal_strength, hal_radius, knots=None) -> FilmFitParams:
        p = FilmFitParams()
        E, nb, K = self.n_emulsions, self.n_bins, self.n_knots
        p.n_emulsions, p.n_bins, p.n_knots = E, nb, K
        np.ctypeslib.as_array(p.sens)[:E, :nb] = np.asarray(sens.detach().cpu().numpy(), np.float32)
        np.ctypeslib.as_array(p.d_max)[:E] = np.asarray(d_max.detach().cpu().numpy(), np.float32)
        g = np.ctypeslib.as_array(p.gamma)
        g[:] = 1.0
        g[:E] = np.asarray(gamma.detach().cpu().numpy(), np.float32)
        if K:
            if knots is None:
                raise ValueError("this stock's curve has knots: pass `knots` (E x n_knots)")
            np.ctypeslib.as_array(p.knot_frac)[:E, :K] = np.asarray(knots.detach().cpu().numpy(), np.float32)
        p.hal_strength = float(hal_strength)
        p.hal_radius = float(hal_radius)
        return p

    def state_like(self, rgb: torch.Tensor) -> torch.Tensor:
        """Device buffer for the forward state that a backward pass reuses."""
        n, h, w, _ = rgb.shape
        floats = C.c_size_t()
        check(lib().imops_film_fit_state_floats(C.byref(self.model), n, w, h, C.byref(floats)))
        return torch.empty(floats.value, dtype=torch.float32, device=rgb.device)

    def forward_raw(self, rgb: torch.Tensor, p: FilmFitParams, state: torch.Tensor | None = None) -> torch.Tensor:
        _check_images(rgb)
        n, h, w, _ = rgb.shape
        out = torch.empty_like(rgb)
        self._after_torch(rgb.device)
        check(lib().imops_film_fit_forward(self.ctx.handle, C.byref(self.model), C.byref(p), rgb.data_ptr(),
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
        check(lib().imops_film_fit_backward(self.ctx.handle, C.byref(self.model), C.byref(p), rgb.data_ptr(),
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
        check(lib().imops_film_fit_l2_step(self.ctx.handle, C.byref(self.model), C.byref(p), rgb.data_ptr(),
                                           target.data_ptr(), n, w, h, C.byref(g), C.byref(loss)))
        return loss.value, g

    def __call__(self, rgb, sens, d_max, gamma, hal_strength, hal_radius, knots=None) -> torch.Tensor:
        return FilmFit.apply(rgb, sens, d_max, gamma, hal_strength, hal_radius, knots, self)


def _check_images(rgb: torch.Tensor) -> None:
    if not (rgb.is_cuda and rgb.dtype == torch.float32 and rgb.dim() == 4 and rgb.shape[-1] == 3
            and rgb.is_contiguous()):
        raise ValueError("film fit images must be a contiguous float32 CUDA tensor of shape (N, H, W, 3)")


class FilmFit(torch.autograd.Function):
    """out = film model(rgb; sens, d_max, gamma, hal_strength, hal_radius[, knots]); gradients
    w.r.t. the parameters."""

    @staticmethod
    def forward(fctx, rgb, sens, d_max, gamma, hal_strength, hal_radius, knots, model):
        p = model.pack(sens, d_max, gamma, hal_strength, hal_radius, knots)
        fctx.model, fctx.packed = model, p
        fctx.has_knots = knots is not None
        # keep the forward's exposure planes for the backward when gradients are wanted
        state = model.state_like(rgb) if any(fctx.needs_input_grad[1:7]) else None
        out = model.forward_raw(rgb, p, state)
        fctx.save_for_backward(rgb, state)
        fctx.like = (sens, d_max, gamma, hal_strength, hal_radius, knots)
        return out

    @staticmethod
    def backward(fctx, grad_out):
        rgb, state = fctx.saved_tensors
        g = fctx.model.unpack(fctx.model.backward_raw(rgb, grad_out, fctx.packed, state))
        sens, d_max, gamma, hs, hr, knots = fctx.like
        as_like = lambda t, like: t.to(dtype=like.dtype, device=like.device).reshape(like.shape)
        return (None, as_like(g["sens"], sens), as_like(g["d_max"], d_max), as_like(g["gamma"], gamma),
                as_like(g["hal_strength"], hs), as_like(g["hal_radius"], hr),
                as_like(g["knots"], knots) if fctx.has_knots else None, None)


# lower bounds of the parameters after every optimiser step
LOWER = {"sens": 0.0, "d_max": 0.05, "gamma": 0.05, "hal_strength": 0.0, "hal_radius": 0.05, "knots": 0.0}


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
    """The fitted parameters as one float32 vector (in `model.names` order: sens E x bins,
    d_max E, gamma E, hal_strength, hal_radius, knots E x n_knots) packed in place into one
    imops_film_fit_params, and the gradient struct read back as a vector of the same layout."""

    def __init__(self, model: FilmModel, init: dict):
        E, nb, K = model.n_emulsions, model.n_bins, model.n_knots
        self.names = model.names
        self.theta = np.concatenate([np.asarray(init[k].detach().cpu().numpy(), np.float32).ravel()
                                     for k in self.names])
        self.lower = np.concatenate([np.full(int(np.prod(init[k].shape)), LOWER[k]) for k in self.names]
                                    ).astype(np.float32)
        self.p = model.pack(**{k: init[k] for k in self.names})
        self.g = FilmFitParams()
        # (struct view, offset into the vector, length) of every field
        self._views = []
        off = 0
        for name, pv, gv, shape in (
                ("sens", np.ctypeslib.as_array(self.p.sens)[:E, :nb], np.ctypeslib.as_array(self.g.sens)[:E, :nb],
                 (E, nb)),
                ("d_max", np.ctypeslib.as_array(self.p.d_max)[:E], np.ctypeslib.as_array(self.g.d_max)[:E], (E,)),
                ("gamma", np.ctypeslib.as_array(self.p.gamma)[:E], np.ctypeslib.as_array(self.g.gamma)[:E], (E,)),
                ("hal_strength", None, None, ()), ("hal_radius", None, None, ()),
                ("knots", np.ctypeslib.as_array(self.p.knot_frac)[:E, :K],
                 np.ctypeslib.as_array(self.g.knot_frac)[:E, :K], (E, K))):
            if name not in self.names:
                continue
            size = int(np.prod(shape))
            self._views.append((name, pv, gv, shape, off, size))
            off += size
        self.grad = np.zeros_like(self.theta)

    def pack(self) -> FilmFitParams:
        t = self.theta
        for name, pv, _, shape, off, size in self._views:
            if pv is None:
                setattr(self.p, name, float(t[off]))
            else:
                pv[...] = t[off:off + size].reshape(shape)
        return self.p

    def read_grad(self) -> np.ndarray:
        g = self.grad
        for name, _, gv, shape, off, size in self._views:
            if gv is None:
                g[off] = getattr(self.g, name)
            else:
                g[off:off + size] = gv.ravel()
        return g

    def as_dict(self) -> dict:
        t = self.theta
        out = {}
        for name, _, _, shape, off, size in self._views:
            out[name] = torch.from_numpy(t[off:off + size].reshape(shape).copy())
        return out


def fit(model: FilmModel, rgb: torch.Tensor, target: torch.Tensor, init: dict, steps: int = 200, lr: float = 0.01,
        log_every: int = 0, fused: bool = True) -> dict:
    """Adam on the L2 loss between the model output and `target`, the parameters clamped
    to LOWER after each step; returns the fitted parameters, the loss history (the loss
    at each step's parameters before its update) and the median wall time of one
    iteration (ms; forward, adjoint and optimiser step).

    fused: each iteration is one imops_film_fit_l2_step (loss and gradient in one device
    pass, one synchronisation) followed by Adam on the host on the flat parameter vector;
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
    params = {k: init[k].detach().clone().requires_grad_(True) for k in model.names}
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
