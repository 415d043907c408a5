"""Differentiable film model: GPU forward against the FP64 oracle, and the
hand-written adjoint against FP64 central finite differences of the oracle and
against the oracle's FP64 reverse pass.

The model is film::process with grain, DIR inhibition and adjacency off and the
stock's sensitivities, characteristic curves and halation as parameters
(oracle: `film_fit`, the oracle's own film pipeline with those switches, and
`film_fit_gradient`, its f64 reverse pass, pinned against central differences in
tests/test_oracle_film_fit_grad_cpu.py).  Two models: ColorNeg200 on the
reference's 16-bin grid (E x 16 sensitivities, d_max, gamma, halation), and the
seeded synthetic stock on the 81-bin grid with its 5-knot curves (3 x 81
sensitivities, 3 x 5 knot fractions, d_max, gamma, halation).
Bars: forward max abs error <= 2e-5 against the FP64 oracle; gradients within
1e-3 relative of FP64 central differences (h = 1e-3) of the oracle's L2 loss and
of the oracle's reverse pass, per entry (against max(|ref|, 1% of the largest
entry)) and as a vector.
"""
import numpy as np
import pytest
import torch

import imops_amd as im
import oracle
from imops_amd import STOCKS
from imops_amd.film_fit import FilmModel, fit

pytestmark = pytest.mark.gpu

PITCH = 36000.0 / 6016  # 35 mm frame at 6016 px: halation sigma 13.4 px, radius 41
STOCK_NAMES = {v: k for k, v in STOCKS.items()}


def _images(n, w, h, seed):
    g = oracle.absolute_gray(1.0, 200)
    return np.stack([(im.synth_scene_rgb(w, h, seed + i) * g).astype(np.float32) for i in range(n)])


def _kw(model):
    m = model.model
    return dict(stock=STOCK_NAMES[m.stock], pitch_um=m.pitch_um, bins=m.spectral_bins, stock_seed=m.stock_seed)


def _np(params):
    return {k: (v.double().numpy() if v.dim() else float(v)) for k, v in params.items()}


def _oracle_forward(model, imgs, params):
    p = _np(params)
    return np.stack([oracle.film_fit(x, p, **_kw(model)) for x in imgs])


def _oracle_loss(model, imgs, params, target):
    out = _oracle_forward(model, imgs, params)
    return float(np.sum((out.astype(np.float64) - target.astype(np.float64)) ** 2))


def _oracle_gradient(model, imgs, params, gouts):
    """Sum over the images of the oracle's f64 reverse pass for dL/dout = gouts."""
    p = _np(params)
    total = None
    for x, gy in zip(imgs, gouts):
        g, _ = oracle.film_fit_gradient(x, gy.astype(np.float32), p, **_kw(model))
        total = g if total is None else {k: total[k] + g[k] for k in g}
    return total


def _perturbed(nom, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = {k: v.clone() for k, v in nom.items()}
    p["sens"] = (p["sens"] * torch.tensor(1 + 0.15 * scale * rng.uniform(-1, 1, p["sens"].shape), dtype=torch.float32))
    p["d_max"] = p["d_max"] + 0.2 * scale * torch.tensor(rng.uniform(-1, 1, p["d_max"].shape), dtype=torch.float32)
    p["gamma"] = p["gamma"] * torch.tensor(1 + 0.1 * scale * rng.uniform(-1, 1, p["gamma"].shape), dtype=torch.float32)
    p["hal_strength"] = torch.tensor(1.0 + 0.3 * scale, dtype=torch.float32)
    p["hal_radius"] = torch.tensor(1.0 - 0.05 * scale, dtype=torch.float32)
    if "knots" in p:
        k = p["knots"] + 0.03 * scale * torch.tensor(rng.uniform(-1, 1, p["knots"].shape), dtype=torch.float32)
        p["knots"] = k.clamp(0.001, 0.999)
    return p


def _vector(model, g):
    """A gradient (struct or dict) as one f64 vector in the model's parameter order."""
    d = model.unpack(g) if not isinstance(g, dict) else g
    return np.concatenate([np.ravel(np.asarray(d[k], np.float64)) for k in model.names])


def _compare(name, got, want, bar_entry=1e-3, bar_vec=1e-3):
    floor = 1e-2 * np.abs(want).max()
    rel = np.abs(got - want) / np.maximum(np.abs(want), floor)
    vec = np.linalg.norm(got - want) / np.linalg.norm(want)
    print(f"{name}: worst entry {rel.max():.2e} (index {int(np.argmax(rel))}), vector {vec:.2e}", flush=True)
    assert rel.max() <= bar_entry and vec <= bar_vec


@pytest.fixture(scope="module")
def model():
    return FilmModel(im.Context(0), pitch_um=PITCH)


@pytest.fixture(scope="module")
def model81():
    """The C4 parameterisation: synthetic stock (seed 7), 81 bins, 5-knot curves."""
    return FilmModel(im.Context(0), "Synthetic", pitch_um=PITCH, spectral_bins=81, stock_seed=7)


@pytest.mark.parametrize("which", ["model", "model81"])
@pytest.mark.parametrize("scale", [0.0, 1.0])
def test_forward_matches_fp64_oracle(request, which, scale):
    m = request.getfixturevalue(which)
    imgs = _images(2, 96, 88, 3)
    p = _perturbed(m.nominal(), 11, scale)
    got = m(torch.from_numpy(imgs).cuda(), **p).cpu().numpy()
    want = _oracle_forward(m, imgs, p)
    err = np.abs(got - want).max()
    print(f"{which}: forward max abs err {err:.3g}", flush=True)
    assert err <= 2e-5


def test_model_shapes(model, model81):
    nom = model81.nominal()
    assert model81.n_bins == 81 and model81.n_knots == 5 and model81.n_emulsions == 3
    assert tuple(nom["sens"].shape) == (3, 81) and tuple(nom["knots"].shape) == (3, 5)
    pos = model81.knot_positions()
    assert pos.shape == (3, 5) and (np.diff(pos, axis=1) > 0).all()
    assert "knots" not in model.nominal() and model.n_knots == 0
    want = oracle.film_fit_nominal("Synthetic", 81, 7)
    np.testing.assert_allclose(nom["knots"].numpy(), want["knots"], rtol=1e-6)
    np.testing.assert_allclose(nom["sens"].numpy(), want["sens"], rtol=1e-6, atol=1e-12)


def test_gradients_match_fp64_central_differences(model):
    E = model.n_emulsions
    imgs = _images(2, 96, 88, 5)
    hidden = _perturbed(model.nominal(), 21, 1.0)
    target = _oracle_forward(model, imgs, hidden)
    theta = _perturbed(model.nominal(), 33, 0.4)

    x = torch.from_numpy(imgs).cuda()
    t = torch.from_numpy(target).cuda()
    params = {k: v.clone().requires_grad_(True) for k, v in theta.items()}
    loss = ((model(x, **params) - t) ** 2).sum()
    loss.backward()

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
        fds.append((_oracle_loss(model, imgs, plus, target) - _oracle_loss(model, imgs, minus, target)) / (2 * h))
        adj.append(float(params[name].grad[idx]))
    adj, fds = np.array(adj), np.array(fds)
    # The oracle rounds a few per-stock scalars to f32 as the reference does (e.g. the
    # halation reflectance), which puts ~1e-3 absolute noise on a difference quotient; entries
    # are therefore compared relative to max(|fd|, 1% of the largest gradient).
    for (name, idx), g, fd in zip(checks, adj, fds):
        print(f"{name}{list(idx)}: adjoint {g:.6g} fd {fd:.6g}", flush=True)
    _compare("adjoint vs central differences", adj, fds)


@pytest.mark.parametrize("which", ["model", "model81"])
def test_reverse_pass_matches_the_fp64_reverse_pass(request, which):
    """Every parameter's gradient (all sensitivities on the grid, knots, curve, halation)
    for one output gradient, GPU against the oracle's f64 reverse pass."""
    m = request.getfixturevalue(which)
    imgs = _images(2, 96, 88, 25)
    theta = _perturbed(m.nominal(), 27, 0.5)
    target = _oracle_forward(m, imgs, _perturbed(m.nominal(), 29, 1.0))
    gout = 2.0 * (_oracle_forward(m, imgs, theta).astype(np.float64) - target)
    x = torch.from_numpy(imgs).cuda()
    got = _vector(m, m.backward_raw(x, torch.from_numpy(gout.astype(np.float32)).cuda(), m.pack(**theta)))
    want = _vector(m, _oracle_gradient(m, imgs, theta, gout))
    _compare(f"{which}: GPU reverse vs FP64 reverse", got, want, bar_entry=1e-3, bar_vec=2e-4)


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


@pytest.mark.parametrize("which", ["model", "model81"])
def test_fused_and_autograd_fits_follow_one_trajectory(request, which):
    """The fused loop (one imops_film_fit_l2_step per iteration, host Adam) and the
    autograd loop (forward, torch loss, FilmFit backward, torch.optim.Adam) take the same
    steps: their losses and final parameters agree to the f32 evaluation of the output."""
    m = request.getfixturevalue(which)
    imgs = _images(2, 96, 88, 17)
    hidden = _perturbed(m.nominal(), 43, 1.0)
    x = torch.from_numpy(imgs).cuda()
    with torch.no_grad():
        target = m(x, **hidden)
    a = fit(m, x, target, m.nominal(), steps=15, lr=0.01, fused=True)
    b = fit(m, x, target, m.nominal(), steps=15, lr=0.01, fused=False)
    la, lb = np.array(a["loss"]), np.array(b["loss"])
    print(f"{which}: loss {la[0]:.4g} -> {la[-1]:.4g}, rel diff {np.abs(la - lb).max() / lb.max():.2e}", flush=True)
    assert la[-1] < la[0]
    np.testing.assert_allclose(la, lb, rtol=1e-4, atol=1e-6 * lb.max())
    for k in a["params"]:
        np.testing.assert_allclose(a["params"][k].numpy(), b["params"][k].numpy(), rtol=1e-4, atol=1e-5)


@pytest.mark.parametrize("which", ["model", "model81"])
def test_l2_step_matches_forward_and_backward(request, which):
    """imops_film_fit_l2_step returns the L2 loss of the forward's output against the
    target and the backward's gradient for dL/dout = 2 (out - target); the step evaluates
    the output in f32 inside the reverse kernel, so both agree to f32 rounding.  Two images
    of different content, plus a shape whose pixel count is not a multiple of the block."""
    m = request.getfixturevalue(which)
    for n, w, h in ((2, 96, 88), (1, 131, 77)):
        imgs = _images(n, w, h, 21)
        x = torch.from_numpy(imgs).cuda()
        p_hidden = m.pack(**_perturbed(m.nominal(), 51, 1.0))
        target = m.forward_raw(x, p_hidden)
        p = m.pack(**_perturbed(m.nominal(), 53, 0.5))
        out = m.forward_raw(x, p)
        want_loss = float(((out.double() - target.double()) ** 2).sum())
        want_g = _vector(m, m.backward_raw(x, (2.0 * (out - target)).contiguous(), p))
        loss, g = m.l2_step(x, target, p)
        got_g = _vector(m, g)
        rel = np.abs(got_g - want_g).max() / np.abs(want_g).max()
        print(f"{which} {n}x{w}x{h}: loss {loss:.8g} vs {want_loss:.8g}, gradient max rel {rel:.2e}", flush=True)
        assert abs(loss - want_loss) <= 1e-5 * want_loss
        assert rel <= 1e-4
        assert np.linalg.norm(got_g - want_g) <= 1e-5 * np.linalg.norm(want_g)


def test_batch_and_saved_state_match_single_images(model):
    """One call over N images (every stage one launch for the whole batch) gives each image
    the bits of a call on that image alone; a backward pass that reads the forward's saved
    state gives the gradients of one that recomputes it, and both equal the sum of the
    single-image gradients (f64 partial sums, so to a few f32 ulps)."""
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
    g_re = _vector(model, model.backward_raw(x, gout, p))
    g_st = _vector(model, model.backward_raw(x, gout, p, state))
    g_sum = sum(_vector(model, model.backward_raw(x[i:i + 1].contiguous(), gout[i:i + 1].contiguous(), p))
                for i in range(3))
    scale = np.abs(g_re).max()
    print(f"recompute vs state {np.abs(g_re - g_st).max() / scale:.2e}, "
          f"batch vs single sum {np.abs(g_re - g_sum).max() / scale:.2e}", flush=True)
    assert np.allclose(g_st, g_re, rtol=1e-6, atol=1e-7 * scale)
    assert np.allclose(g_sum, g_re, rtol=1e-5, atol=1e-6 * scale)


def test_gradients_at_the_c4_shape_match_fp64_central_differences(model):
    """The 16-bin model at the adjoint workload's shape: 8 ACEScg crops of 512x512.  Per
    parameter (d_max, gamma, halation, four sensitivity bins per emulsion) and along two
    random directions over every parameter, the GPU adjoint's gradient of the L2 loss
    against FP64 central differences (h = 1e-3) of the FP64 oracle's loss; the fused L2
    step against the same differences."""
    E = model.n_emulsions
    print("film fit: gradients at 8x512x512", flush=True)
    imgs = _images(8, 512, 512, 60)
    hidden = _perturbed(model.nominal(), 61, 1.0)
    target = _oracle_forward(model, imgs, hidden)
    theta = _perturbed(model.nominal(), 63, 0.4)

    x = torch.from_numpy(imgs).cuda()
    t = torch.from_numpy(target).cuda()
    params = {k: v.clone().requires_grad_(True) for k, v in theta.items()}
    loss = ((model(x, **params) - t) ** 2).sum()
    loss.backward()
    grads = {k: v.grad.detach().double() for k, v in params.items()}

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
        fds.append((_oracle_loss(model, imgs, moved(d, 1, h), target) -
                    _oracle_loss(model, imgs, moved(d, -1, h), target)) / (2 * h))
        adj.append(float(grads[name][idx]))
    adj, fds = np.array(adj), np.array(fds)
    for (name, idx), g, fd in zip(checks, adj, fds):
        print(f"{name}{list(idx)}: adjoint {g:.6g} fd {fd:.6g}", flush=True)
    _compare("adjoint vs central differences", adj, fds)
    # directional derivatives along random directions over every parameter (sensitivities scaled
    # to their magnitude so each coordinate moves by a comparable relative amount)
    rng = np.random.default_rng(65)
    for k in range(2):
        d = {name: torch.from_numpy(rng.uniform(-1, 1, tuple(v.shape))) * (v.double().abs() + 0.1)
             for name, v in theta.items()}
        fd = (_oracle_loss(model, imgs, moved(d, 1, h), target) -
              _oracle_loss(model, imgs, moved(d, -1, h), target)) / (2 * h)
        g = sum(float((grads[name] * d[name]).sum()) for name in d)
        r = abs(g - fd) / abs(fd)
        print(f"direction {k}: adjoint {g:.6g} fd {fd:.6g} rel {r:.2e}", flush=True)
        assert r <= 1e-3
    # the fused L2 step at the same point: its loss against the oracle's, its gradient
    # against the same central differences
    loss, g = model.l2_step(x, t, model.pack(**theta))
    want = _oracle_loss(model, imgs, theta, target)
    gv = model.unpack(g)
    fused = np.array([float(np.asarray(gv[name])[idx]) for name, idx in checks])
    print(f"fused step: loss {loss:.8g} vs oracle {want:.8g}", flush=True)
    assert abs(loss - want) <= 1e-4 * want
    _compare("fused step vs central differences", fused, fds)


def test_c4_parameterisation_gradients_at_the_c4_shape(model81):
    """The C4 workload's model and shape: the synthetic stock on the 81-bin grid with
    5-knot curves (parameters 3 x 81 sensitivities, 3 x 5 knot fractions, d_max, gamma,
    halation radius and strength), 8 ACEScg crops of 512x512.  The GPU adjoint's L2 gradient
    (autograd) and the fused L2 step's, against (1) the oracle's FP64 reverse pass, every
    parameter, and (2) FP64 central differences (h = 1e-3) of the oracle's loss for the
    halation pair, the curve, a middle knot and the sensitivity peak of each emulsion, and
    along two random directions over every parameter."""
    m = model81
    E = m.n_emulsions
    print("film fit: 81-bin knot model at 8x512x512", flush=True)
    imgs = _images(8, 512, 512, 70)
    target = _oracle_forward(m, imgs, _perturbed(m.nominal(), 71, 1.0))
    theta = _perturbed(m.nominal(), 73, 0.4)

    x = torch.from_numpy(imgs).cuda()
    t = torch.from_numpy(target).cuda()
    params = {k: v.clone().requires_grad_(True) for k, v in theta.items()}
    loss = ((m(x, **params) - t) ** 2).sum()
    loss.backward()
    got = _vector(m, {k: v.grad for k, v in params.items()})
    fused_loss, g = m.l2_step(x, t, m.pack(**theta))
    fused = _vector(m, g)

    # (1) the FP64 reverse pass of the oracle's own L2 gradient
    out64 = _oracle_forward(m, imgs, theta).astype(np.float64)
    want = _vector(m, _oracle_gradient(m, imgs, theta, 2.0 * (out64 - target)))
    _compare("adjoint vs FP64 reverse pass", got, want)
    _compare("fused step vs FP64 reverse pass", fused, want)
    want_loss = float(np.sum((out64 - target) ** 2))
    print(f"fused loss {fused_loss:.8g} vs oracle {want_loss:.8g}", flush=True)
    assert abs(fused_loss - want_loss) <= 1e-4 * want_loss

    # (2) central differences of the oracle's loss
    def moved(direction, sign, h):
        return {k: (v.double() + sign * h * direction[k]).float() for k, v in theta.items()}

    h = 1e-3
    peaks = [int(torch.argmax(theta["sens"][e])) for e in range(E)]
    checks = [("hal_strength", ()), ("hal_radius", ())] + [("d_max", (e,)) for e in range(E)] + \
             [("gamma", (e,)) for e in range(E)] + [("knots", (e, 2)) for e in range(E)] + \
             [("sens", (e, peaks[e])) for e in range(E)]
    offsets, off = {}, 0
    for k in m.names:
        offsets[k] = off
        off += int(np.prod(theta[k].shape))
    fds, sel = [], []
    for name, idx in checks:
        d = {k: torch.zeros_like(v, dtype=torch.float64) for k, v in theta.items()}
        d[name][idx] = 1.0
        fds.append((_oracle_loss(m, imgs, moved(d, 1, h), target) -
                    _oracle_loss(m, imgs, moved(d, -1, h), target)) / (2 * h))
        flat = offsets[name] + (int(np.ravel_multi_index(idx, tuple(theta[name].shape))) if idx else 0)
        sel.append(flat)
    fds = np.array(fds)
    for (name, idx), fd, i in zip(checks, fds, sel):
        print(f"{name}{list(idx)}: adjoint {got[i]:.6g} fused {fused[i]:.6g} fp64 reverse {want[i]:.6g} fd {fd:.6g}",
              flush=True)
    _compare("adjoint vs central differences", got[sel], fds)
    _compare("fused step vs central differences", fused[sel], fds)
    rng = np.random.default_rng(75)
    for k in range(2):
        d = {name: torch.from_numpy(rng.uniform(-1, 1, tuple(v.shape))) * (v.double().abs() + 0.1)
             for name, v in theta.items()}
        fd = (_oracle_loss(m, imgs, moved(d, 1, h), target) - _oracle_loss(m, imgs, moved(d, -1, h), target)) / (2 * h)
        dv = _vector(m, {name: v.numpy() for name, v in d.items()})
        ga, gf = float(got @ dv), float(fused @ dv)
        mag = float(np.abs(want * dv).sum())
        print(f"direction {k}: adjoint {ga:.6g} fused {gf:.6g} fd {fd:.6g} "
              f"rel {abs(ga - fd) / abs(fd):.2e}, error / magnitude {abs(ga - fd) / mag:.2e}", flush=True)
        # relative to the sum of the terms' magnitudes (the groups' contributions cancel)
        assert abs(ga - fd) <= 1e-3 * mag and abs(gf - fd) <= 1e-3 * mag
