"""Pins the FP64 guide oracle (oracle/guide.py) with checks that do not retype
it: central finite differences of the loss for the analytic gradient,
quadrature over the sphere for the vMF density, a sampled histogram against
that density, Adam's closed-form first step, and bfloat16 rounding of
representative values."""
import math

import numpy as np

from oracle import guide as G


def _rand_problem(n, seed):
    rng = np.random.default_rng(seed)
    params = rng.normal(0, 0.3, G.n_params())
    x = rng.random((n, 3))
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    wo = rng.normal(size=(n, 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = rng.random(n) * 2
    return params, G.encode(x, nrm, wo), d, w


def test_parameter_count_and_layout():
    assert G.n_params() == 15080
    p = np.arange(G.n_params(), dtype=np.float64)
    layers = G.unpack(p)
    assert [W.shape for W, _ in layers] == [(64, 64), (64, 64), (64, 64), (64, 40)]
    assert layers[1][0][0, 0] == 64 * 64 + 64 and layers[3][1][-1] == G.n_params() - 1


def test_output_gradient_matches_finite_differences():
    params, enc, d, w = _rand_problem(16, 1)
    out, _ = G.forward(params, enc)
    g = G.output_grad(out, d, w)

    def L(o):
        mu, kappa, pi = G.lobes(o)
        return -(w * np.log(G.mixture_pdf(mu, kappa, pi, d))).sum() / len(w)

    h = 1e-6
    fd = np.zeros_like(out)
    for i in range(out.shape[0]):
        for j in range(out.shape[1]):
            op, om = out.copy(), out.copy()
            op[i, j] += h
            om[i, j] -= h
            fd[i, j] = (L(op) - L(om)) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-9)


def test_analytic_gradient_matches_finite_differences():
    params, enc, d, w = _rand_problem(24, 2)
    g = G.loss_grad(params, enc, d, w)
    rng = np.random.default_rng(3)
    idx = np.concatenate([rng.choice(G.n_params(), 150, replace=False), [0, 4095, 4096 * 3 + 191, G.n_params() - 1]])
    h = 1e-6
    for j in idx:
        pp, pm = params.copy(), params.copy()
        pp[j] += h
        pm[j] -= h
        fd = (G.loss(pp, enc, d, w) - G.loss(pm, enc, d, w)) / (2 * h)
        assert abs(g[j] - fd) <= 1e-5 * max(1.0, abs(fd)) + 1e-8, (j, g[j], fd)


def _sphere_grid(mu, nt=4000, npf=256):
    """Directions and solid-angle weights of a (theta, phi) grid about mu,
    dense near theta = 0 (theta = pi s^2)."""
    s = (np.arange(nt) + 0.5) / nt
    th = math.pi * s * s
    dth = 2 * math.pi * s / nt
    ph = (np.arange(npf) + 0.5) / npf * 2 * math.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    t, b = G.frame(np.asarray(mu, np.float64))
    dirs = (np.cos(T)[..., None] * mu + np.sin(T)[..., None] *
            (np.cos(P)[..., None] * t + np.sin(P)[..., None] * b)).reshape(-1, 3)
    dw = (np.sin(T) * dth[:, None] * (2 * math.pi / npf)).reshape(-1)
    return dirs, dw, T.reshape(-1)


def test_vmf_pdf_integrates_to_one():
    mu = np.array([0.3, -0.5, 0.81])
    mu /= np.linalg.norm(mu)
    dirs, dw, _ = _sphere_grid(mu)
    for kappa in (0.05, 1.0, 30.0, 1000.0):
        assert abs((G.vmf_pdf(mu, kappa, dirs) * dw).sum() - 1.0) < 2e-3, kappa


def test_vmf_sampling_follows_pdf():
    rng = np.random.default_rng(5)
    mu = np.array([-0.2, 0.9, 0.1])
    mu /= np.linalg.norm(mu)
    for kappa in (0.5, 8.0, 200.0):
        n = 200_000
        d = G.vmf_sample(mu, kappa, rng.random(n), rng.random(n))
        assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-9)
        ang = np.arccos(np.clip(d @ mu, -1, 1))
        edges = np.quantile(ang, np.linspace(0, 1, 21))
        obs = np.histogram(ang, edges)[0]
        grid, dw, th = _sphere_grid(mu, nt=6000, npf=8)
        p = G.vmf_pdf(mu, kappa, grid) * dw
        exp = np.array([p[(th >= lo) & (th < hi)].sum() for lo, hi in zip(edges[:-1], edges[1:])]) * n
        z = np.abs(obs - exp) / np.sqrt(exp)
        assert z.max() < 6, (kappa, z.max())
        # the azimuth about mu is uniform
        t, b = G.frame(mu)
        phi = np.arctan2(d @ b, d @ t)
        h = np.histogram(phi, 16, (-math.pi, math.pi))[0]
        assert np.all(np.abs(h - n / 16) < 6 * math.sqrt(n / 16))


def test_mixture_weights_and_directions():
    out = np.zeros((1, 40))
    out[0, 4::5] = np.log(np.arange(1, 9))
    out[0, 0:3] = [0, 0, 2]
    mu, kappa, pi = G.lobes(out)
    np.testing.assert_allclose(pi[0], np.arange(1, 9) / 36.0, rtol=1e-12)
    np.testing.assert_allclose(mu[0, 0], [0, 0, 1])
    np.testing.assert_allclose(kappa[0], 1.0)


def test_adam_first_step_moves_each_parameter_by_lr():
    rng = np.random.default_rng(6)
    p = rng.normal(size=1000)
    g = rng.normal(size=1000) * 10 ** rng.uniform(-3, 2, 1000)
    p2, m, v = G.adam_step(p, g, np.zeros(1000), np.zeros(1000), 1, lr=1e-3)
    # step 1: the bias-corrected moments are g and g^2, so |delta| = lr |g| / (|g| + eps)
    np.testing.assert_allclose(np.abs(p2 - p), 1e-3 * np.abs(g) / (np.abs(g) + 1e-8), rtol=1e-9)
    assert np.all(np.sign(p - p2) == np.sign(g))


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 1e-20, 65504.0], np.float32)
    r = G.bf16(v)
    # ties round to even (1 + 2^-8 -> 1), above the tie rounds up, 65504 -> 2^16
    np.testing.assert_array_equal(r[[0, 1, 2, 3, 5]], [1.0, 1.0, 1.0 + 2 ** -7, -2.5, 65536.0])
    assert abs(r[4] - 1e-20) / 1e-20 < 2 ** -8
