"""Pins the FP64 guide oracle (oracle/guide.py) by mathematics: finite differences for
the gradient, sphere quadrature for the vMF densities, the closed-form mean cosine of
the vMF sampler, closed forms of the encoding and of Adam's first step.  The guide has
no counterpart in the reference, so there are no golden vectors for it."""
import numpy as np
import pytest

from oracle import guide as G


def _batch(rng, n):
    X = rng.uniform(-1, 1, (n, G.N_IN))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = rng.uniform(0.1, 2.0, n)
    return X, d, w


def _params(rng, scale=0.3):
    return rng.normal(0, scale, G.N_PARAMS)


def test_parameter_count_matches_header():
    assert G.N_PARAMS == 14056


@pytest.mark.parametrize("emulate", [False, True])
def test_gradient_matches_central_differences(emulate):
    rng = np.random.default_rng(7)
    X, d, w = _batch(rng, 24)
    p = _params(rng)
    _, g = G.loss_and_grad(p, X, d, w, emulate_bf16=False)
    # every parameter kind: first/last weight rows of each layer, every bias vector,
    # the output block of directions, concentrations and logits
    offs, off = [], 0
    for o, i in G.SHAPES:
        offs += [off, off + 1, off + o * i - 1, off + o * i, off + o * i + o - 1]
        off += o * i + o
    W4 = off - 40 * 64 - 40
    offs += [W4 + 64 * r + 5 for r in (0, 2, 23, 24, 31, 32, 39)]
    offs += [off - 40 + j for j in (0, 1, 23, 24, 31, 32, 39)]
    h = 1e-6
    for k in offs:
        e = np.zeros_like(p)
        e[k] = h
        num = (G.loss(p + e, X, d, w) - G.loss(p - e, X, d, w)) / (2 * h)
        assert abs(num - g[k]) <= 1e-6 + 1e-5 * abs(num), (k, num, g[k])
    if emulate:  # the bf16 variant differs from fp64 only by rounding
        _, gq = G.loss_and_grad(p, X, d, w, emulate_bf16=True)
        cos = gq @ g / np.linalg.norm(gq) / np.linalg.norm(g)
        assert cos > 0.99


def test_planted_sign_error_is_caught():
    """A dropped term in the kappa derivative fails the finite-difference check."""
    rng = np.random.default_rng(3)
    X, d, w = _batch(rng, 16)
    p = _params(rng)
    _, g = G.loss_and_grad(p, X, d, w)
    k = G.N_PARAMS - 40 + 24  # bias of the first concentration output
    e = np.zeros_like(p)
    e[k] = 1e-6
    num = (G.loss(p + e, X, d, w) - G.loss(p - e, X, d, w)) / 2e-6
    # restate d log f / d kappa without its normaliser term: must disagree
    O, _ = G.forward(p, X)
    mu, kappa, wt = G.decode(O)
    f = G.vmf_pdf(mu, kappa, d[:, None, :])
    q = np.sum(wt * f, -1)
    c = np.sum(mu * d[:, None, :], -1)
    gk = wt[:, 0] * f[:, 0] / q * (1.0 / kappa[:, 0] + c[:, 0] - 1.0)
    sig = 1.0 / (1.0 + np.exp(-O[:, 24]))
    wrong = np.sum(-w / w.sum() * gk * sig)
    assert abs(num - g[k]) < 1e-6 + 1e-5 * abs(num)
    assert abs(num - wrong) > 1e-4


def _sphere_grid(n=400):
    th = (np.arange(n) + 0.5) / n * np.pi
    ph = (np.arange(2 * n) + 0.5) / (2 * n) * 2 * np.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1).reshape(-1, 3)
    dA = (np.sin(T) * (np.pi / n) * (2 * np.pi / (2 * n))).ravel()
    return d, dA


@pytest.mark.parametrize("kappa", [0.01, 1.0, 10.0, 100.0])
def test_vmf_pdf_integrates_to_one(kappa):
    d, dA = _sphere_grid(800 if kappa > 50 else 300)
    mu = np.array([0.3, 0.5, -0.812403840463596])
    mu /= np.linalg.norm(mu)
    v = G.vmf_pdf(mu[None, :], np.float64(kappa), d)
    assert abs(np.sum(v * dA) - 1.0) < 2e-3


def test_mixture_pdf_integrates_to_one():
    rng = np.random.default_rng(5)
    O = rng.normal(0, 1.5, (1, 40))
    mix = G.decode(O)
    d, dA = _sphere_grid(500)
    v = G.mixture_pdf(tuple(np.repeat(a, d.shape[0], 0) for a in mix), d)
    assert abs(np.sum(v * dA) - 1.0) < 2e-3
    assert np.allclose(mix[2].sum(1), 1.0)


@pytest.mark.parametrize("kappa", [0.5, 5.0, 50.0])
def test_vmf_sample_mean_cosine_closed_form(kappa):
    rng = np.random.default_rng(11)
    n = 200000
    mu = np.array([[-0.48, 0.6, 0.64]])
    mu /= np.linalg.norm(mu)
    mu = np.repeat(mu, n, 0)
    s = G.vmf_sample(mu, kappa, rng.uniform(size=n), rng.uniform(size=n))
    assert np.allclose(np.linalg.norm(s, axis=1), 1.0, atol=1e-9)
    mean_cos = np.mean(np.sum(s * mu, 1))
    expect = 1.0 / np.tanh(kappa) - 1.0 / kappa
    assert abs(mean_cos - expect) < 4e-3
    # the transverse components average to zero (isotropy around mu)
    t = s - np.sum(s * mu, 1, keepdims=True) * mu
    assert np.linalg.norm(t.mean(0)) < 4e-3


def test_encoding_closed_forms():
    f = G.encode([[0.25, 0.5, 0.0]], [[0.0, 1.0, -1.0]], [[1.0, 0.0, 0.0]])[0]
    assert f.shape == (48,)
    # x = 1/4: sin(pi/2)=1, cos=0; sin(pi)=0, cos=-1; then 2 pi, 4 pi -> (0, 1)
    assert np.allclose(f[:8], [1, 0, 0, -1, 0, 1, 0, 1], atol=1e-12)
    assert np.allclose(f[8:10], [0, -1], atol=1e-12)  # y = 1/2, k = 0
    assert np.allclose(f[16:24], [0, 1] * 4, atol=1e-12)  # z = 0
    # n.y = 1 -> u = 1: bins at 1/8, 3/8, 5/8, 7/8 -> exp(-0.5 (4 (1 - c))^2)
    c = (np.arange(4) + 0.5) / 4
    assert np.allclose(f[28:32], np.exp(-0.5 * (4 * (1 - c)) ** 2))
    assert np.allclose(f[32:36], np.exp(-0.5 * (4 * (0 - c)) ** 2))  # n.z = -1 -> u = 0


def test_device_rounded_encoding_is_the_encoding():
    """The encoding with the device's float32 argument roundings stays within a few 1e-6
    of the exact one (and meets the same closed forms), and it does differ: the roundings
    are what it adds."""
    rng = np.random.default_rng(4)
    n = 20000
    p01 = rng.uniform(size=(n, 3)).astype(np.float32)
    nrm = rng.normal(size=(n, 3))
    nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    wo = rng.normal(size=(n, 3))
    wo = (wo / np.linalg.norm(wo, axis=1, keepdims=True)).astype(np.float32)
    exact, dev = G.encode(p01, nrm, wo), G.encode_device_rounding(p01, nrm, wo)
    err = np.abs(dev - exact)
    assert err.max() < 2e-5 and err.mean() < 1e-6
    assert err.max() > 0.0
    f = G.encode_device_rounding([[0.25, 0.5, 0.0]], [[0.0, 1.0, -1.0]], [[1.0, 0.0, 0.0]])[0]
    assert np.allclose(f, G.encode([[0.25, 0.5, 0.0]], [[0.0, 1.0, -1.0]], [[1.0, 0.0, 0.0]])[0], atol=2e-6)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(2)
    p = rng.normal(size=1000).astype(np.float32)
    g = rng.normal(size=1000).astype(np.float32) * 1e-2
    z = np.zeros_like(p)
    p1, m1, m2 = G.adam_step(p, z, z, g, 1, lr=1e-2)
    # bias-corrected first step: m/bc1 = g, v/bc2 = g^2 -> -lr g / (|g| + eps)
    assert np.allclose(p1 - p, -1e-2 * g / (np.abs(g) + 1e-8), rtol=1e-4, atol=1e-7)
    assert np.allclose(m1, 0.1 * g, rtol=1e-5)
