"""FP64 restatement of the guiding network (encoding, MLP, vMF mixture, loss, gradient,
Adam) used to check the HIP kernels of toaster_amd/csrc/th_guide.{h,hip}.

TEST INFRASTRUCTURE ONLY: only tests/ import this module; the shipped library never does.

The guide has no counterpart in the reference (its diffuse vertices sample the cosine
lobe only, shaders/pathtrace_spheres.wgsl:849-857), so there are no golden vectors.
Each function is pinned by mathematics instead (tests/test_guide_oracle.py):
- `loss_and_grad`: central finite differences of `loss` on every parameter kind;
- `vmf_pdf` / `mixture_pdf`: quadrature over the sphere integrates to 1;
- `vmf_sample`: the closed-form mean cosine coth(kappa) - 1/kappa and unit length;
- `encode`: closed forms (sin/cos at dyadic points, Gaussian bin values);
- `adam_step`: the first step equals -lr * g / (|g| + eps') for Adam's bias correction.

Parameter layout (14,056 floats): W1 [64][48], b1 [64], W2 [64][64], b2, W3 [64][64], b3,
W4 [40][64], b4 [40]; weights are stored out-major (row n holds the inputs of unit n).
"""
import numpy as np

N_IN, N_HID, N_OUT, LOBES = 48, 64, 40, 8
SHAPES = [(64, 48), (64, 64), (64, 64), (40, 64)]
N_PARAMS = sum(o * i + o for o, i in SHAPES)
assert N_PARAMS == 14056


def split(params):
    """Views (W, b) per layer into the flat parameter vector."""
    out, off = [], 0
    for o, i in SHAPES:
        W = params[off:off + o * i].reshape(o, i)
        off += o * i
        b = params[off:off + o]
        off += o
        out.append((W, b))
    return out


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    finite = (u & 0x7F800000) != 0x7F800000
    r = np.where(finite, (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000, u & 0xFFFF0000)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def encode(p01, n, wo):
    """Features (B x 48): per axis of p01, sin/cos(2 pi 2^k x) for k = 0..3 (24), then
    each of the 6 direction components mapped to u = (c + 1)/2 and spread over 4 Gaussian
    bins at (j + 1/2)/4 with sigma 1/4 (24)."""
    p01, n, wo = (np.atleast_2d(np.asarray(a, np.float64)) for a in (p01, n, wo))
    cols = []
    for a in range(3):
        for k in range(4):
            ang = 2.0 * np.pi * (2.0 ** k) * p01[:, a]
            cols += [np.sin(ang), np.cos(ang)]
    dc = np.concatenate([n, wo], 1)
    for a in range(6):
        u = 0.5 * (dc[:, a] + 1.0)
        for j in range(4):
            d = (u - (j + 0.5) * 0.25) * 4.0
            cols.append(np.exp(-0.5 * d * d))
    return np.stack(cols, 1)


def encode_device_rounding(p01, n, wo):
    """`encode` with the float32 roundings of the device's arguments (inputs taken as
    float32): the angle f32(f32(6.2831855 * 2^k) * x), handed to sine and cosine in
    revolutions as f32(angle * f32(1 / (2 pi))); the bin offset d = f32(f32(u - c) * 4) of
    u = f32(0.5 * f32(c + 1)), and the exponent f32(f32(-0.5 * d) * d) handed to the base-2
    exponential as f32(exponent * f32(log2 e)).  Each function is then evaluated exactly,
    so the features differ from `encode` only by those argument roundings (a few 1e-6) --
    which matters where a feature lies that close to a bfloat16 rounding boundary."""
    f = np.float32
    p01, n, wo = (np.atleast_2d(np.asarray(a, f)) for a in (p01, n, wo))
    cols = []
    for a in range(3):
        for k in range(4):
            ang = f(f(6.2831855) * f(2.0 ** k)) * p01[:, a]
            rev = (ang * f(0.15915494)).astype(np.float64)
            cols += [np.sin(2.0 * np.pi * rev), np.cos(2.0 * np.pi * rev)]
    dc = np.concatenate([n, wo], 1)
    for a in range(6):
        u = f(0.5) * (dc[:, a] + f(1.0))
        for j in range(4):
            d = (u - f((j + 0.5) * 0.25)) * f(4.0)
            x = (f(-0.5) * d) * d
            cols.append(np.exp2((x * f(1.4426950408889634)).astype(np.float64)))
    return np.stack(cols, 1)


def forward(params, X, emulate_bf16=False):
    """Returns (O, cache).  With emulate_bf16 the inputs, weights and hidden activations
    are rounded to bfloat16 as on the device (accumulation stays exact here)."""
    q = bf16 if emulate_bf16 else (lambda v: np.asarray(v, np.float64))
    layers = split(np.asarray(params, np.float64))
    h = q(X)
    acts = [h]
    zs = []
    for li, (W, b) in enumerate(layers):
        z = h @ q(W).T + b
        zs.append(z)
        if li < 3:
            h = q(np.maximum(z, 0.0))
            acts.append(h)
        else:
            h = z
    return h, (acts, zs)


def softplus(x):
    return np.where(x > 15.0, x, np.log1p(np.exp(np.minimum(x, 15.0))))


def decode(O):
    """Mixture (mu [B,8,3], kappa [B,8], w [B,8]) from network outputs [B,40]."""
    O = np.atleast_2d(O)
    r = O[:, :24].reshape(-1, LOBES, 3)
    ln = np.sqrt(np.sum(r * r, -1) + 1e-8)
    mu = r / ln[..., None]
    kappa = np.minimum(softplus(O[:, 24:32]) + 0.01, 1000.0)
    lg = O[:, 32:40]
    e = np.exp(lg - lg.max(1, keepdims=True))
    return mu, kappa, e / e.sum(1, keepdims=True)


def vmf_pdf(mu, kappa, w):
    """kappa / (2 pi (1 - exp(-2 kappa))) exp(kappa (mu.w - 1))."""
    c = np.sum(mu * w, -1)
    return kappa / (2.0 * np.pi * (-np.expm1(-2.0 * kappa))) * np.exp(kappa * (c - 1.0))


def mixture_pdf(mix, w):
    mu, kappa, wt = mix
    return np.sum(wt * vmf_pdf(mu, kappa, np.asarray(w)[:, None, :]), -1)


def loss(params, X, dirs, weights, emulate_bf16=False):
    O, _ = forward(params, X, emulate_bf16)
    q = np.maximum(mixture_pdf(decode(O), dirs), 1e-30)
    return float(-np.sum(weights * np.log(q)) / np.sum(weights))


def loss_and_grad(params, X, dirs, weights, emulate_bf16=False):
    """Loss -sum w log q(dir) / sum w and its gradient by backpropagation."""
    params = np.asarray(params, np.float64)
    O, (acts, zs) = forward(params, X, emulate_bf16)
    mu, kappa, wt = decode(O)
    dirs = np.asarray(dirs, np.float64)
    f = vmf_pdf(mu, kappa, dirs[:, None, :])
    q = np.maximum(np.sum(wt * f, -1), 1e-30)
    S = np.sum(weights)
    L = float(-np.sum(weights * np.log(q)) / S)
    scale = (-weights / S)[:, None]
    g = wt * f / q[:, None]  # responsibilities
    r = O[:, :24].reshape(-1, LOBES, 3)
    ln = np.sqrt(np.sum(r * r, -1) + 1e-8)
    c = np.sum(mu * dirs[:, None, :], -1)
    dr = (dirs[:, None, :] - mu * c[..., None]) * (g * kappa / ln)[..., None]
    e2 = np.exp(-2.0 * kappa)
    dk = g * (1.0 / kappa - 2.0 * e2 / (1.0 - e2) + (c - 1.0))
    sig = np.where(kappa >= 1000.0, 0.0, 1.0 / (1.0 + np.exp(-O[:, 24:32])))
    dO = np.concatenate([dr.reshape(-1, 24), dk * sig, g - wt], 1) * scale
    grads = []
    layers = split(params)
    dz = dO
    for li in range(3, -1, -1):
        W, _ = layers[li]
        Wq = bf16(W) if emulate_bf16 else W
        grads.append((dz.T @ acts[li], dz.sum(0)))
        if li > 0:
            dz = (dz @ Wq) * (acts[li] > 0)
    grads.reverse()
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return L, flat


def adam_step(params, m1, m2, grad, step, lr=1e-2, b1=0.9, b2=0.99, eps=1e-8):
    """One Adam update in float32 arithmetic order of the device kernel; returns new
    (params, m1, m2).  `step` counts from 1."""
    f = np.float32
    params, m1, m2, grad = (np.asarray(a, f) for a in (params, m1, m2, grad))
    a = f(b1) * m1 + (f(1.0) - f(b1)) * grad
    v = f(b2) * m2 + (f(1.0) - f(b2)) * grad * grad
    bc1 = f(1.0) - np.power(f(b1), f(step))
    bc2 = f(1.0) - np.power(f(b2), f(step))
    p = params - f(lr) * (a / bc1) / (np.sqrt(v / bc2) + f(eps))
    return p, a, v


def vmf_sample(mu, kappa, u1, u2):
    """Inversion of the cosine marginal (Jakob 2012) in a basis around mu."""
    w = 1.0 + np.log(u1 + (1.0 - u1) * np.exp(-2.0 * kappa)) / kappa
    w = np.clip(w, -1.0, 1.0)
    r = np.sqrt(np.maximum(0.0, 1.0 - w * w))
    s = np.where(mu[..., 2] >= 0, 1.0, -1.0)
    a = -1.0 / (s + mu[..., 2])
    b = mu[..., 0] * mu[..., 1] * a
    t = np.stack([1.0 + s * mu[..., 0] ** 2 * a, s * b, -s * mu[..., 0]], -1)
    bb = np.stack([b, s + mu[..., 1] ** 2 * a, -mu[..., 1]], -1)
    ph = 2.0 * np.pi * u2
    return (w[..., None] * mu + (r * np.cos(ph))[..., None] * t
            + (r * np.sin(ph))[..., None] * bb)
