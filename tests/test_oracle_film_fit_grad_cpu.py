"""The oracle's FP64 reverse pass of the differentiable film model, pinned against
FP64 central differences of the oracle's own forward.

The reverse pass (`oracle.film_fit_gradient`, oracle/film.cpp) is written from the
forward's equations step by step; the forward is `oracle.film_fit` in f64.  The
checks cover every parameter group on both grids: the reference's 16-bin grid with
the ColorNeg200 crystal LUTs, and the 81-bin grid with the seeded synthetic stock's
5-knot curves (the knot fractions are parameters there).  Small frames at a coarse
pitch keep the blurs short (local radius 1, halation radius 4) so every reflect-101
edge case is inside the frame.  A dropped term (the coloured mask's dM/dD, the
halation bounce's gain, a transmittance of a layer above) or a transposed blur
applied the wrong way round moves at least one entry far past the bar.

Bars: each entry within 5e-4 of its difference quotient, relative to max(|fd|, 1%
of the largest entry), 1e-4 as a vector and 3e-4 along a random direction
(relative to the sum of its terms' magnitudes).  The
forward rounds d_max, gamma, the halation reflectance, gain and blur width to f32
(as the reference stores them); the reverse takes those roundings as the identity,
which bounds the agreement near 1e-4 on the entries they touch.
"""
import numpy as np
import pytest

import imops_amd as im
import oracle

PITCH = 60.0  # um per pixel: halation sigma 1.33 px (radius 4), local scatter radius 1


def _images(n, w, h, seed):
    g = oracle.absolute_gray(1.0, 200)
    return np.stack([(im.synth_scene_rgb(w, h, seed + i) * g).astype(np.float32) for i in range(n)])


def _perturbed(nom, seed, s):
    rng = np.random.default_rng(seed)
    q = {k: (np.array(v, np.float64) if isinstance(v, np.ndarray) else float(v)) for k, v in nom.items()}
    q["sens"] = q["sens"] * (1 + 0.15 * s * rng.uniform(-1, 1, q["sens"].shape))
    q["d_max"] = q["d_max"] + 0.2 * s * rng.uniform(-1, 1, q["d_max"].shape)
    q["gamma"] = q["gamma"] * (1 + 0.1 * s * rng.uniform(-1, 1, q["gamma"].shape))
    q["hal_strength"] = 1.0 + 0.3 * s
    q["hal_radius"] = 1.0 - 0.05 * s
    if "knots" in q:
        q["knots"] = np.clip(q["knots"] + 0.03 * s * rng.uniform(-1, 1, q["knots"].shape), 0.001, 0.999)
    return q


@pytest.mark.parametrize("stock,bins,seed", [("ColorNeg200", 16, 0), ("Synthetic", 81, 7)])
def test_reverse_pass_matches_central_differences(stock, bins, seed):
    kw = dict(stock=stock, pitch_um=PITCH, bins=bins, stock_seed=seed)
    x = _images(2, 40, 32, 5)
    nom = oracle.film_fit_nominal(stock, bins, seed)
    E = len(nom["d_max"])
    hidden, theta = _perturbed(nom, 1, 1.0), _perturbed(nom, 2, 0.4)
    target = np.stack([oracle.film_fit(xi, hidden, **kw) for xi in x]).astype(np.float64)

    def loss(p):
        out = np.stack([oracle.film_fit(xi, p, **kw) for xi in x]).astype(np.float64)
        return float(((out - target) ** 2).sum())

    out = np.stack([oracle.film_fit(xi, theta, **kw) for xi in x])
    grad = None
    for i in range(len(x)):
        g, fwd = oracle.film_fit_gradient(x[i], (2.0 * (out[i] - target[i])).astype(np.float32), theta, **kw)
        # the reverse pass runs the same forward
        assert np.abs(fwd - out[i]).max() <= 1e-6 * np.abs(out[i]).max()
        grad = g if grad is None else {k: grad[k] + g[k] for k in g}

    # steps: 1e-3, and 5e-4 for the sensitivities, whose difference quotients' truncation
    # error at 1e-3 reaches 8e-4 where a sensitivity is near zero, while below 5e-4 the f32
    # rounding of the halation reflectance and gain they feed dominates the quotient
    steps = {"sens": 5e-4}
    h = 1e-3
    checks = [("d_max", (e,)) for e in range(E)] + [("gamma", (e,)) for e in range(E)] + \
             [("hal_strength", ()), ("hal_radius", ())] + \
             [("sens", (e, i)) for e in range(E) for i in (1, bins // 3, bins // 2, bins - 3)]
    if "knots" in nom:
        checks += [("knots", (e, k)) for e in range(E) for k in range(nom["knots"].shape[1])]
    adj, fds = [], []
    for name, idx in checks:
        hs = steps.get(name, h)

        def moved(sign):
            q = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in theta.items()}
            if idx == ():
                q[name] += sign * hs
            else:
                q[name][idx] += sign * hs
            return q
        fds.append((loss(moved(1)) - loss(moved(-1))) / (2 * hs))
        adj.append(float(np.asarray(grad[name])[idx]))
    adj, fds = np.array(adj), np.array(fds)
    floor = 1e-2 * np.abs(fds).max()
    rel = np.abs(adj - fds) / np.maximum(np.abs(fds), floor)
    for (name, idx), a, f, r in zip(checks, adj, fds, rel):
        print(f"{name}{list(idx)}: reverse {a:.6g} fd {f:.6g} rel {r:.2e}")
    vec = np.linalg.norm(adj - fds) / np.linalg.norm(fds)
    print(f"worst entry {rel.max():.2e}, vector {vec:.2e}")
    assert rel.max() <= 5e-4 and vec <= 1e-4
    # and along a random direction over every parameter
    rng = np.random.default_rng(9)
    d = {k: rng.uniform(-1, 1, np.shape(v)) * (np.abs(v) + 0.1) for k, v in theta.items()}
    plus = {k: theta[k] + h * d[k] for k in theta}
    minus = {k: theta[k] - h * d[k] for k in theta}
    fd = (loss(plus) - loss(minus)) / (2 * h)
    dd = sum(float(np.sum(np.asarray(grad[k]) * d[k])) for k in d)
    # relative to the sum of the terms' magnitudes: the groups' contributions cancel
    # (e.g. +5.2 from the curve, -6.1 from the knots), which would inflate a ratio to |fd|
    mag = sum(float(np.sum(np.abs(np.asarray(grad[k]) * d[k]))) for k in d)
    print(f"direction: reverse {dd:.6g} fd {fd:.6g} error / magnitude {abs(dd - fd) / mag:.2e}")
    assert abs(dd - fd) <= 3e-4 * mag


def test_reverse_pass_is_linear_in_the_output_gradient():
    """dL/dparams is linear in dL/dout: the gradient for a + 2b equals grad(a) + 2 grad(b)."""
    kw = dict(stock="Synthetic", pitch_um=PITCH, bins=81, stock_seed=7)
    x = _images(1, 24, 20, 8)[0]
    theta = _perturbed(oracle.film_fit_nominal("Synthetic", 81, 7), 3, 0.5)
    rng = np.random.default_rng(4)
    # multiples of 1/256 below 8 in magnitude: a + 2 b is exact in f32
    a = (np.round(np.clip(rng.normal(size=x.shape), -3, 3) * 256) / 256).astype(np.float32)
    b = (np.round(np.clip(rng.normal(size=x.shape), -3, 3) * 256) / 256).astype(np.float32)
    ga, _ = oracle.film_fit_gradient(x, a, theta, **kw)
    gb, _ = oracle.film_fit_gradient(x, b, theta, **kw)
    gc, _ = oracle.film_fit_gradient(x, (a + 2 * b).astype(np.float32), theta, **kw)
    for k in ga:
        want = np.asarray(ga[k]) + 2 * np.asarray(gb[k])
        np.testing.assert_allclose(np.asarray(gc[k]), want, rtol=1e-9, atol=1e-12 * max(1.0, np.abs(want).max()))
