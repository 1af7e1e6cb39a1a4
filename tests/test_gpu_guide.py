"""Guiding network kernels on the GPU against the FP64 oracle (oracle/guide.py): the
MFMA fragment mapping, forward outputs, loss gradient, the Adam update and a training
run that fits a known directional distribution.

Tolerances: the device rounds inputs, weights, hidden activations and the output
gradient to BF16 and accumulates in FP32; the oracle rounds the same values to BF16
(emulate_bf16) and accumulates exactly.  Outputs: >= 99% of rows agree to
|d| <= 1e-4 (1 + |ref|) and every row to 0.1 (1 + |ref|) -- a hidden activation whose
exact value lies within FP32 accumulation error of a BF16 rounding boundary rounds to
the neighbouring BF16 value (one BF16 ulp, up to 2^-6 of a unit of magnitude 4), which
the following layers carry to the outputs.  Gradients: cosine >= 0.999 and relative
norm error <= 3e-2 (the BF16 output gradient)."""
import numpy as np
import pytest

from oracle import guide as G
from toaster_amd.guide import Guide, bf16_selftest, mfma32_selftest, mfma_selftest

pytestmark = pytest.mark.gpu


def test_mfma_fragment_mapping(gpu_ctx):
    rng = np.random.default_rng(0)
    a = rng.normal(size=(16, 32)).astype(np.float32)
    b = rng.normal(size=(32, 16)).astype(np.float32)
    print("case mfma selftest", flush=True)
    c = mfma_selftest(gpu_ctx, a, b)
    ref = G.bf16(a) @ G.bf16(b)
    assert np.allclose(c, ref, rtol=1e-5, atol=1e-5)


def test_mfma32_fragment_mapping(gpu_ctx):
    """The 32x32x16 mapping the inference kernel relies on (lane s + 32 h holds
    A[s][8h..8h+7] and B[8h..8h+7][s]; accumulator r of that lane is D[8 (r/4) + 4 h + r % 4][s])."""
    rng = np.random.default_rng(1)
    a = rng.normal(size=(32, 16)).astype(np.float32)
    b = rng.normal(size=(16, 32)).astype(np.float32)
    print("case mfma32 selftest", flush=True)
    c = mfma32_selftest(gpu_ctx, a, b)
    ref = G.bf16(a) @ G.bf16(b)
    assert np.allclose(c, ref, rtol=1e-5, atol=1e-5)


def test_bf16_rounding_matches_oracle(gpu_ctx):
    """The device's float -> BF16 conversion (the gfx950 instruction every network kernel
    uses) gives the oracle's integer round-to-nearest-even (oracle/guide.py bf16) for all
    2^32 float patterns except NaNs, denormals and infinities included."""
    print("case bf16 selftest over 2^32 patterns", flush=True)
    bad, _ = bf16_selftest(gpu_ctx)
    assert bad == 0


def _inputs(rng, n):
    p01 = rng.uniform(size=(n, 3))
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    wo = rng.normal(size=(n, 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    return G.encode(p01, nrm, wo).astype(np.float32)


@pytest.mark.parametrize("n", [1, 37, 4096, 70000])
def test_forward_matches_oracle(gpu_ctx, n):
    rng = np.random.default_rng(n)
    g = Guide(gpu_ctx, seed=3)
    p = rng.normal(0, 0.25, G.N_PARAMS).astype(np.float32)
    g.set_params(p)
    assert np.array_equal(g.get_params(), p)
    X = _inputs(rng, n)
    print(f"case forward n={n}", flush=True)
    O = g.forward(X)
    ref, _ = G.forward(p, X, emulate_bf16=True)
    err = np.abs(O - ref)
    row_ok = np.all(err <= 1e-4 * (1 + np.abs(ref)), axis=1)
    print(f"  rows within 1e-4: {row_ok.mean():.5f}, max err {err.max():.3g}")
    assert row_ok.mean() >= 0.99
    assert np.all(err <= 0.1 * (1 + np.abs(ref))), float(err.max())
    assert float(err.mean()) < 1e-3
    g.close()


@pytest.mark.parametrize("n", [1, 37, 4096, 70000])
def test_fused_inference_matches_oracle(gpu_ctx, n):
    """The guided renderer's inference kernel (encoding on the device, then the four
    layers) on host-supplied vertices against the oracle's encoding and forward pass.
    The oracle encodes with the device's float32 argument roundings
    (encode_device_rounding); the transcendentals themselves are the hardware's, within a
    few ulp, so a feature rarely rounds to the neighbouring BF16 value -- the forward
    test's row tolerance covers it.  The kernel stores its outputs as binary16, so each
    output may also differ by one binary16 ulp (2^-10 of its magnitude) from the oracle's
    exact value."""
    rng = np.random.default_rng(500 + n)
    g = Guide(gpu_ctx, seed=3)
    p = rng.normal(0, 0.25, G.N_PARAMS).astype(np.float32)
    g.set_params(p)
    bmin = np.array([-4.0, -0.5, -2.5], np.float32)
    bmax = np.array([4.0, 1.6, 2.0], np.float32)
    pos = (bmin + rng.uniform(size=(n, 3)) * (bmax - bmin)).astype(np.float32)
    nrm = rng.normal(size=(n, 3))
    nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    wo = rng.normal(size=(n, 3))
    wo = (wo / np.linalg.norm(wo, axis=1, keepdims=True)).astype(np.float32)
    print(f"case fused inference n={n}", flush=True)
    O = g.infer_vertices(pos, nrm, wo, bmin, bmax)
    # the device's normalisation: (p - bmin) * f32(1 / (bmax - bmin))
    scale = (np.float32(1.0) / (bmax - bmin)).astype(np.float32)
    p01 = ((pos - bmin) * scale).astype(np.float32)
    ref, _ = G.forward(p, G.encode_device_rounding(p01, nrm, wo), emulate_bf16=True)
    err = np.abs(O - ref)
    row_ok = np.all(err <= 1e-4 * (1 + np.abs(ref)) + 2.0 ** -10 * np.abs(ref), axis=1)
    print(f"  rows within 1e-4: {row_ok.mean():.5f}, max err {err.max():.3g}")
    assert row_ok.mean() >= 0.99
    assert np.all(err <= 0.1 * (1 + np.abs(ref))), float(err.max())
    assert float(err.mean()) < 1e-3
    g.close()


@pytest.mark.parametrize("n", [64, 5000, 262144])
def test_gradient_and_adam_match_oracle(gpu_ctx, n):
    rng = np.random.default_rng(100 + n)
    g = Guide(gpu_ctx, seed=5)
    p = rng.normal(0, 0.2, G.N_PARAMS).astype(np.float32)
    g.set_params(p)
    X = _inputs(rng, n)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = rng.uniform(0.0, 2.0, n)
    print(f"case gradient n={n}", flush=True)
    loss, grad = g.train_batch(X, d, w, update=False)
    ref_loss, ref_grad = G.loss_and_grad(p, X, d.astype(np.float32), w.astype(np.float32),
                                         emulate_bf16=True)
    assert abs(loss - ref_loss) <= 2e-3 * (1 + abs(ref_loss))
    cos = float(grad @ ref_grad / np.linalg.norm(grad) / np.linalg.norm(ref_grad))
    rel = float(np.linalg.norm(grad - ref_grad) / np.linalg.norm(ref_grad))
    print(f"  loss {loss:.6f} ref {ref_loss:.6f} cos {cos:.6f} rel {rel:.4f}")
    assert cos >= 0.999 and rel <= 3e-2
    assert np.array_equal(g.get_params(), p)  # gradient only: no update
    # one Adam step: the device update equals the oracle's Adam on the device gradient
    g.set_hyper(lr=1e-2)
    _, g1 = g.train_batch(X, d, w, update=True)
    assert np.allclose(g1, grad, rtol=1e-5, atol=1e-9)
    z = np.zeros_like(p)
    p1, _, _ = G.adam_step(p, z, z, g1, 1, lr=1e-2)
    assert np.allclose(g.get_params(), p1, rtol=1e-6, atol=1e-7)
    g.close()


def test_training_fits_a_directional_distribution(gpu_ctx):
    """Directions drawn from a vMF whose mean depends on position: Adam at lr 1e-2 on
    batches of 2^14 drives the loss (cross-entropy in nats) from the uniform sphere's
    log(4 pi) = 2.53 to near the target's entropy (kappa 20: -0.30)."""
    rng = np.random.default_rng(9)
    g = Guide(gpu_ctx, seed=1)
    g.set_hyper(lr=1e-2)

    def batch(n):
        p01 = rng.uniform(size=(n, 3))
        ang = 2 * np.pi * p01[:, 0]
        mu = np.stack([np.cos(ang), 0.3 + 0 * ang, np.sin(ang)], 1)
        mu /= np.linalg.norm(mu, axis=1, keepdims=True)
        d = G.vmf_sample(mu, 20.0, rng.uniform(size=n), rng.uniform(size=n))
        nrm = np.tile([0.0, 1.0, 0.0], (n, 1))
        X = G.encode(p01, nrm, nrm).astype(np.float32)
        return X, d, np.ones(n)

    X, d, w = batch(1 << 14)
    first, _ = g.train_batch(X, d, w, update=False)
    losses = []
    print("case training 300 steps", flush=True)
    for it in range(300):
        X, d, w = batch(1 << 14)
        l, _ = g.train_batch(X, d, w, update=True)
        losses.append(l)
    last = float(np.mean(losses[-20:]))
    print(f"  loss {first:.3f} -> {last:.3f}")
    assert first > 1.5
    assert last < 0.5
    g.close()
