"""The learned guiding distribution: an MLP (48 -> 64 -> 64 -> 64 -> 40, BF16 on MFMA)
giving an 8-lobe von Mises-Fisher mixture over incident directions at diffuse vertices.
It has no counterpart in the reference, whose diffuse vertices sample the cosine lobe
only (shaders/pathtrace_spheres.wgsl:849-857)."""
import ctypes as C

import numpy as np

from . import _abi as A
from . import renderer
from .renderer import GpuContext, GpuScene


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Guide:
    """Owns one network (FP32 master weights, Adam moments, BF16 copies) on a context."""

    def __init__(self, ctx: GpuContext, seed=1):
        self.ctx = ctx
        h = C.c_void_p()
        A.check(A.lib().th_guide_create(ctx.handle, seed, C.byref(h)))
        self._h = h
        renderer._dependents.add(self)
        ctx.adopt(self)

    @property
    def handle(self):
        return self._h

    def get_params(self):
        out = np.empty(A.TH_GUIDE_PARAMS, np.float32)
        A.check(A.lib().th_guide_get_params(self._h, _ptr(out), out.size))
        return out

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        A.check(A.lib().th_guide_set_params(self._h, _ptr(p), p.size))

    def set_hyper(self, lr=1e-2, beta1=0.9, beta2=0.99, eps=1e-8):
        A.check(A.lib().th_guide_set_hyper(self._h, lr, beta1, beta2, eps))

    def forward(self, features):
        f = np.ascontiguousarray(features, np.float32)
        assert f.ndim == 2 and f.shape[1] == A.TH_GUIDE_FEATURES
        out = np.empty((f.shape[0], A.TH_GUIDE_OUTPUTS), np.float32)
        A.check(A.lib().th_guide_forward(self._h, _ptr(f), f.shape[0], _ptr(out)))
        return out

    def infer_vertices(self, pos, nrm, wo, bmin, bmax):
        """Outputs (count x 40) of the renderer's fused inference kernel for diffuse
        vertices (position, normal, outgoing direction; position normalised by the box
        [bmin, bmax] on the device)."""
        v = np.ascontiguousarray(np.concatenate([np.asarray(a, np.float32).reshape(-1, 3)
                                                 for a in (pos, nrm, wo)], 1), np.float32)
        lo = np.ascontiguousarray(bmin, np.float32)
        hi = np.ascontiguousarray(bmax, np.float32)
        assert lo.shape == (3,) and hi.shape == (3,)
        out = np.empty((v.shape[0], A.TH_GUIDE_OUTPUTS), np.float32)
        A.check(A.lib().th_guide_infer_vertices(self._h, _ptr(v), v.shape[0], _ptr(lo), _ptr(hi), _ptr(out)))
        return out

    def train_batch(self, features, dirs, weights, update=True):
        """One Adam step (or the gradient only); returns (loss, gradient)."""
        f = np.ascontiguousarray(features, np.float32)
        d = np.ascontiguousarray(dirs, np.float32)
        w = np.ascontiguousarray(weights, np.float32)
        n = f.shape[0]
        assert f.shape == (n, A.TH_GUIDE_FEATURES) and d.shape == (n, 3) and w.shape == (n,)
        loss = np.zeros(1, np.float32)
        grad = np.empty(A.TH_GUIDE_PARAMS, np.float32)
        A.check(A.lib().th_guide_train_batch(self._h, _ptr(f), _ptr(d), _ptr(w), n,
                                             1 if update else 0, _ptr(loss), _ptr(grad)))
        return float(loss[0]), grad

    def close(self):
        if self._h:
            A.lib().th_guide_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mfma32_selftest(ctx: GpuContext, a, b):
    """C = A (32x16) B (16x32) through one v_mfma_f32_32x32x16_bf16 with the inference
    kernel's fragment mapping (inputs rounded to BF16)."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    assert a.shape == (32, 16) and b.shape == (16, 32)
    c = np.empty((32, 32), np.float32)
    A.check(A.lib().th_guide_mfma32_selftest(ctx.handle, _ptr(a), _ptr(b), _ptr(c)))
    return c


def mfma_selftest(ctx: GpuContext, a, b):
    """C = A (16x32) B (32x16) through one v_mfma_f32_16x16x32_bf16 with this library's
    fragment mapping (inputs rounded to BF16)."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    assert a.shape == (16, 32) and b.shape == (32, 16)
    c = np.empty((16, 16), np.float32)
    A.check(A.lib().th_guide_mfma_selftest(ctx.handle, _ptr(a), _ptr(b), _ptr(c)))
    return c


def bf16_selftest(ctx: GpuContext):
    """(non-NaN mismatches, NaN mismatches) of the device's float -> BF16 rounding against
    integer round-to-nearest-even over all 2^32 float patterns."""
    m, mn = C.c_uint64(), C.c_uint64()
    A.check(A.lib().th_bf16_selftest(ctx.handle, C.byref(m), C.byref(mn)))
    return m.value, mn.value


# Defaults measured on the indirect-lit terrain (1280x720, 262,144 triangles): guide
# probability 0.8, environment light by the continuation alone at diffuse vertices, Adam
# steps of 2^15 samples, a warm-up of 768 steps at 16 steps per pass (48 passes) with
# alpha ramped over the first 512, then one step every pass.
GUIDED_DEFAULTS = dict(alpha=0.8, env_nee=False, batch=1 << 15, warmup_steps=768, warmup_rate=16,
                       alpha_ramp=512, train_every=1)


def render_guided(scene: GpuScene, guide: Guide, samples, alpha=0.8, train=True, steps_per_pass=1,
                  batch=1 << 15, progressive=True, env_nee=True, warmup_steps=768, warmup_rate=16,
                  alpha_ramp=512, train_every=1, alpha_from_start=False, time_network=False):
    """`samples` passes of one sample per pixel with the guide (see th_render_guided);
    with progressive=True the running average and frame seed continue like
    GpuScene.render_progressive.  env_nee=False drops the environment sample at every
    diffuse vertex (TH_GUIDED_NO_ENV_NEE); the default keeps the reference's estimator,
    which alpha = 0 then reproduces.  alpha_from_start=True uses the network from the
    first pass even before it has taken a step (TH_GUIDED_ALPHA_FROM_START: an untrained
    network with train=False).  time_network=True times the network's launches
    (TH_GUIDED_TIME_NETWORK).  Returns the call's statistics as a dict."""
    p = scene.params(samples=samples)
    if not progressive:
        p.accumulated_samples = 0
    o = A.GuidedOptions(alpha=alpha, train=1 if train else 0, steps_per_pass=steps_per_pass,
                        batch=batch, flags=((0 if env_nee else A.TH_GUIDED_NO_ENV_NEE)
                                            | (A.TH_GUIDED_ALPHA_FROM_START if alpha_from_start else 0)
                                            | (A.TH_GUIDED_TIME_NETWORK if time_network else 0)),
                        warmup_steps=warmup_steps, warmup_rate=warmup_rate, alpha_ramp=alpha_ramp,
                        train_every=train_every)
    st = A.GuidedStats()
    A.check(A.lib().th_render_guided(scene.ctx.handle, scene._h, guide.handle, C.byref(p),
                                     C.byref(scene.packed.camera), C.byref(o), C.byref(st)))
    if progressive:
        scene.accumulated_samples = p.accumulated_samples + samples
        scene.frame_index += 1
    return {"diffuse_vertices": st.diffuse_vertices, "records": st.records,
            "train_steps": st.train_steps, "last_loss": st.last_loss, "inferred_rows": st.inferred_rows,
            "trained_samples": st.trained_samples, "infer_ms": st.infer_ms, "train_ms": st.train_ms}


def rel_rmse(image, reference):
    """sqrt(mean over pixels and RGB channels of (I - R)^2 / (R^2 + 0.01))."""
    i = np.asarray(image, np.float64)[..., :3]
    r = np.asarray(reference, np.float64)[..., :3]
    return float(np.sqrt(np.mean((i - r) ** 2 / (r * r + 0.01))))
