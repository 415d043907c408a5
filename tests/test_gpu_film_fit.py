"""Differentiable film model: GPU forward against the FP64 oracle, and the
hand-written adjoint against FP64 central finite differences of the oracle.

The model is film::process with grain, DIR inhibition and adjacency off and the
stock's sensitivities, characteristic curves and halation as parameters
(oracle: `film_fit`, the oracle's own film pipeline with those switches).
Bars: forward max abs error <= 2e-5 against the FP64 oracle; gradients within
1e-3 relative of FP64 central differences (h = 1e-3) of the oracle's L2 loss,
per entry (against max(|fd|, 1% of the largest entry)) and as a vector.
"""
import numpy as np
import pytest
import torch

import imops_amd as im
import oracle
from imops_amd.film_fit import FilmModel, fit

pytestmark = pytest.mark.gpu

PITCH = 36000.0 / 6016  # 35 mm frame at 6016 px: halation sigma 13.4 px, radius 41


def _images(n, w, h, seed):
    g = oracle.absolute_gray(1.0, 200)
    return np.stack([(im.synth_scene_rgb(w, h, seed + i) * g).astype(np.float32) for i in range(n)])


def _flat(params, E):
    v = np.zeros(oracle.FIT_NPARAMS)
    v[0] = E
    s = np.zeros((8, 16))
    s[:E] = params["sens"].double().numpy()
    v[1:129] = s.ravel()
    v[129:129 + E] = params["d_max"].double().numpy()
    v[137:137 + E] = params["gamma"].double().numpy()
    v[137 + E:145] = 1.0
    v[145] = float(params["hal_strength"])
    v[146] = float(params["hal_radius"])
    return v


def _oracle_forward(imgs, params, E):
    v = _flat(params, E)
    return np.stack([oracle.film_fit(x, v, pitch_um=PITCH) for x in imgs])


def _perturbed(nom, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = {k: v.clone() for k, v in nom.items()}
    p["sens"] = (p["sens"] * torch.tensor(1 + 0.15 * scale * rng.uniform(-1, 1, p["sens"].shape), dtype=torch.float32))
    p["d_max"] = p["d_max"] + 0.2 * scale * torch.tensor(rng.uniform(-1, 1, p["d_max"].shape), dtype=torch.float32)
    p["gamma"] = p["gamma"] * torch.tensor(1 + 0.1 * scale * rng.uniform(-1, 1, p["gamma"].shape), dtype=torch.float32)
    p["hal_strength"] = torch.tensor(1.0 + 0.3 * scale, dtype=torch.float32)
    p["hal_radius"] = torch.tensor(1.0 - 0.05 * scale, dtype=torch.float32)
    return p


@pytest.fixture(scope="module")
def model():
    return FilmModel(im.Context(0), pitch_um=PITCH)


@pytest.mark.parametrize("scale", [0.0, 1.0])
def test_forward_matches_fp64_oracle(model, scale):
    imgs = _images(2, 96, 88, 3)
    p = _perturbed(model.nominal(), 11, scale)
    got = model(torch.from_numpy(imgs).cuda(), **p).cpu().numpy()
    want = _oracle_forward(imgs, p, model.n_emulsions)
    err = np.abs(got - want).max()
    print(f"forward max abs err {err:.3g}", flush=True)
    assert err <= 2e-5


def test_gradients_match_fp64_central_differences(model):
    E = model.n_emulsions
    imgs = _images(2, 96, 88, 5)
    hidden = _perturbed(model.nominal(), 21, 1.0)
    target = _oracle_forward(imgs, hidden, E)
    theta = _perturbed(model.nominal(), 33, 0.4)

    x = torch.from_numpy(imgs).cuda()
    t = torch.from_numpy(target).cuda()
    params = {k: v.clone().requires_grad_(True) for k, v in theta.items()}
    loss = ((model(x, **params) - t) ** 2).sum()
    loss.backward()

    def oracle_loss(p):
        out = _oracle_forward(imgs, p, E)
        return float(np.sum((out.astype(np.float64) - target.astype(np.float64)) ** 2))

    h = 1e-3
    checks = [("d_max", (e,)) for e in range(E)] + [("gamma", (e,)) for e in range(E)] + \
             [("hal_strength", ()), ("hal_radius", ())] + \
             [("sens", (e, i)) for e in range(E) for i in (2, 5, 8, 12)]
    adj, fds = [], []
    for name, idx in checks:
        plus = {k: v.clone() for k, v in theta.items()}
        minus = {k: v.clone() for k, v in theta.items()}
        plus[name][idx] += h
        minus[name][idx] -= h
        fds.append((oracle_loss(plus) - oracle_loss(minus)) / (2 * h))
        adj.append(float(params[name].grad[idx]))
    adj, fds = np.array(adj), np.array(fds)
    # The oracle rounds a few per-stock scalars to f32 as the reference does (e.g. the
    # halation reflectance), which puts ~1e-3 absolute noise on a difference quotient; entries
    # are therefore compared relative to max(|fd|, 1% of the largest gradient).
    floor = 1e-2 * np.abs(fds).max()
    rel = np.abs(adj - fds) / np.maximum(np.abs(fds), floor)
    for (name, idx), g, fd, r in zip(checks, adj, fds, rel):
        print(f"{name}{list(idx)}: adjoint {g:.6g} fd {fd:.6g} rel {r:.2e}", flush=True)
    vec = np.linalg.norm(adj - fds) / np.linalg.norm(fds)
    print(f"worst entry {rel.max():.2e}, vector relative error {vec:.2e}", flush=True)
    assert rel.max() <= 1e-3 and vec <= 1e-3


@pytest.mark.parametrize("fused", [True, False])
def test_fit_reduces_loss(model, fused):
    imgs = _images(2, 96, 88, 7)
    hidden = _perturbed(model.nominal(), 41, 1.0)
    x = torch.from_numpy(imgs).cuda()
    with torch.no_grad():
        target = model(x, **hidden)
    res = fit(model, x, target, model.nominal(), steps=30, lr=0.01, fused=fused)
    print(f"fused={fused}: loss {res['loss'][0]:.4g} -> {res['loss'][-1]:.4g}, {res['ms_per_iter']:.2f} ms/iter",
          flush=True)
    assert res["loss"][-1] < 0.5 * res["loss"][0]


def test_fused_and_autograd_fits_follow_one_trajectory(model):
    """The fused loop (one imops_film_fit_l2_step per iteration, host Adam) and the
    autograd loop (forward, torch loss, FilmFit backward, torch.optim.Adam) take the same
    steps: their losses and final parameters agree to the f32 evaluation of the output."""
    imgs = _images(2, 96, 88, 17)
    hidden = _perturbed(model.nominal(), 43, 1.0)
    x = torch.from_numpy(imgs).cuda()
    with torch.no_grad():
        target = model(x, **hidden)
    a = fit(model, x, target, model.nominal(), steps=15, lr=0.01, fused=True)
    b = fit(model, x, target, model.nominal(), steps=15, lr=0.01, fused=False)
    la, lb = np.array(a["loss"]), np.array(b["loss"])
    print(f"loss rel diff {np.abs(la - lb).max() / lb.max():.2e}", flush=True)
    np.testing.assert_allclose(la, lb, rtol=1e-4, atol=1e-6 * lb.max())
    for k in a["params"]:
        np.testing.assert_allclose(a["params"][k].numpy(), b["params"][k].numpy(), rtol=1e-4, atol=1e-5)


def test_l2_step_matches_forward_and_backward(model):
    """imops_film_fit_l2_step returns the L2 loss of the forward's output against the
    target and the backward's gradient for dL/dout = 2 (out - target); the step evaluates
    the output in f32 inside the reverse kernel, so both agree to f32 rounding.  Two images
    of different content, plus a shape whose pixel count is not a multiple of the block."""
    E = model.n_emulsions
    for n, w, h in ((2, 96, 88), (1, 131, 77)):
        imgs = _images(n, w, h, 21)
        x = torch.from_numpy(imgs).cuda()
        p_hidden = model.pack(**_perturbed(model.nominal(), 51, 1.0))
        target = model.forward_raw(x, p_hidden)
        p = model.pack(**_perturbed(model.nominal(), 53, 0.5))
        out = model.forward_raw(x, p)
        want_loss = float(((out.double() - target.double()) ** 2).sum())
        want_g = _grad_vector(model.backward_raw(x, (2.0 * (out - target)).contiguous(), p), E)
        loss, g = model.l2_step(x, target, p)
        got_g = _grad_vector(g, E)
        rel = np.abs(got_g - want_g).max() / np.abs(want_g).max()
        print(f"{n}x{w}x{h}: loss {loss:.8g} vs {want_loss:.8g}, gradient max rel {rel:.2e}", flush=True)
        assert abs(loss - want_loss) <= 1e-5 * want_loss
        assert rel <= 1e-4
        assert np.linalg.norm(got_g - want_g) <= 1e-5 * np.linalg.norm(want_g)


def _grad_vector(g, E):
    return np.concatenate([np.ctypeslib.as_array(g.sens)[:E].ravel(), np.ctypeslib.as_array(g.d_max)[:E],
                           np.ctypeslib.as_array(g.gamma)[:E], [g.hal_strength, g.hal_radius]]).astype(np.float64)


def test_batch_and_saved_state_match_single_images(model):
    """One call over N images (every stage one launch for the whole batch) gives each image
    the bits of a call on that image alone; a backward pass that reads the forward's saved
    state gives the gradients of one that recomputes it, and both equal the sum of the
    single-image gradients (f64 partial sums, so to a few f32 ulps)."""
    E = model.n_emulsions
    imgs = _images(3, 96, 88, 9)
    p = model.pack(**_perturbed(model.nominal(), 5, 0.5))
    x = torch.from_numpy(imgs).cuda()
    whole = model.forward_raw(x, p).cpu().numpy()
    for i in range(3):
        single = model.forward_raw(x[i:i + 1].contiguous(), p).cpu().numpy()
        assert np.array_equal(whole[i].view(np.uint32), single[0].view(np.uint32))
    state = model.state_like(x)
    saved = model.forward_raw(x, p, state).cpu().numpy()
    assert np.array_equal(saved.view(np.uint32), whole.view(np.uint32))
    gout = torch.from_numpy(np.random.default_rng(3).normal(size=imgs.shape).astype(np.float32)).cuda()
    g_re = _grad_vector(model.backward_raw(x, gout, p), E)
    g_st = _grad_vector(model.backward_raw(x, gout, p, state), E)
    g_sum = sum(_grad_vector(model.backward_raw(x[i:i + 1].contiguous(), gout[i:i + 1].contiguous(), p), E)
                for i in range(3))
    scale = np.abs(g_re).max()
    print(f"recompute vs state {np.abs(g_re - g_st).max() / scale:.2e}, "
          f"batch vs single sum {np.abs(g_re - g_sum).max() / scale:.2e}", flush=True)
    assert np.allclose(g_st, g_re, rtol=1e-6, atol=1e-7 * scale)
    assert np.allclose(g_sum, g_re, rtol=1e-5, atol=1e-6 * scale)


def test_gradients_at_the_c4_shape_match_fp64_central_differences(model):
    """The same check at the adjoint workload's shape: 8 ACEScg crops of 512x512.  Per
    parameter (d_max, gamma, halation, four sensitivity bins per emulsion) and along two
    random directions over every parameter, the GPU adjoint's gradient of the L2 loss
    against FP64 central differences (h = 1e-3) of the FP64 oracle's loss."""
    E = model.n_emulsions
    print("film fit: gradients at 8x512x512", flush=True)
    imgs = _images(8, 512, 512, 60)
    hidden = _perturbed(model.nominal(), 61, 1.0)
    target = _oracle_forward(imgs, hidden, E)
    theta = _perturbed(model.nominal(), 63, 0.4)

    x = torch.from_numpy(imgs).cuda()
    t = torch.from_numpy(target).cuda()
    params = {k: v.clone().requires_grad_(True) for k, v in theta.items()}
    loss = ((model(x, **params) - t) ** 2).sum()
    loss.backward()
    grads = {k: v.grad.detach().double() for k, v in params.items()}

    def oracle_loss(p):
        out = _oracle_forward(imgs, p, E)
        return float(np.sum((out.astype(np.float64) - target.astype(np.float64)) ** 2))

    def moved(direction, sign, h):
        return {k: (v.double() + sign * h * direction[k]).float() for k, v in theta.items()}

    h = 1e-3
    checks = [("d_max", (e,)) for e in range(E)] + [("gamma", (e,)) for e in range(E)] + \
             [("hal_strength", ()), ("hal_radius", ())] + \
             [("sens", (e, i)) for e in range(E) for i in (2, 5, 8, 12)]
    adj, fds = [], []
    for name, idx in checks:
        d = {k: torch.zeros_like(v, dtype=torch.float64) for k, v in theta.items()}
        d[name][idx] = 1.0
        fds.append((oracle_loss(moved(d, 1, h)) - oracle_loss(moved(d, -1, h))) / (2 * h))
        adj.append(float(grads[name][idx]))
    adj, fds = np.array(adj), np.array(fds)
    floor = 1e-2 * np.abs(fds).max()
    rel = np.abs(adj - fds) / np.maximum(np.abs(fds), floor)
    for (name, idx), g, fd, r in zip(checks, adj, fds, rel):
        print(f"{name}{list(idx)}: adjoint {g:.6g} fd {fd:.6g} rel {r:.2e}", flush=True)
    vec = np.linalg.norm(adj - fds) / np.linalg.norm(fds)
    print(f"worst entry {rel.max():.2e}, vector relative error {vec:.2e}", flush=True)
    assert rel.max() <= 1e-3 and vec <= 1e-3
    # directional derivatives along random directions over every parameter (sensitivities scaled
    # to their magnitude so each coordinate moves by a comparable relative amount)
    rng = np.random.default_rng(65)
    for k in range(2):
        d = {name: torch.from_numpy(rng.uniform(-1, 1, tuple(v.shape))) * (v.double().abs() + 0.1)
             for name, v in theta.items()}
        fd = (oracle_loss(moved(d, 1, h)) - oracle_loss(moved(d, -1, h))) / (2 * h)
        g = sum(float((grads[name] * d[name]).sum()) for name in d)
        r = abs(g - fd) / abs(fd)
        print(f"direction {k}: adjoint {g:.6g} fd {fd:.6g} rel {r:.2e}", flush=True)
        assert r <= 1e-3
    # the fused L2 step at the same point: its loss against the oracle's, its gradient
    # against the same central differences
    loss, g = model.l2_step(x, t, model.pack(**theta))
    want = oracle_loss(theta)
    gv = {"d_max": np.ctypeslib.as_array(g.d_max)[:E], "gamma": np.ctypeslib.as_array(g.gamma)[:E],
          "hal_strength": np.float32(g.hal_strength), "hal_radius": np.float32(g.hal_radius),
          "sens": np.ctypeslib.as_array(g.sens)[:E]}
    fused = np.array([float(np.asarray(gv[name])[idx]) for name, idx in checks])
    rel = np.abs(fused - fds) / np.maximum(np.abs(fds), floor)
    vec = np.linalg.norm(fused - fds) / np.linalg.norm(fds)
    print(f"fused step: loss {loss:.8g} vs oracle {want:.8g}; worst entry {rel.max():.2e}, vector {vec:.2e}",
          flush=True)
    assert abs(loss - want) <= 1e-4 * want
    assert rel.max() <= 1e-3 and vec <= 1e-3
