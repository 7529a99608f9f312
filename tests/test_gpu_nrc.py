"""GPU parity of the neural-radiance-cache kernels against the FP64 oracle.

  MFMA lane layout  exact products of small-integer matrices
  inference         fp16 / bf16 network vs the oracle with the same 16-bit
                    operand rounding (tolerance: 16-bit activation rounding)
  gradients         per-layer relative error of the summed gradient
  Adam              one step reproduces the bias-corrected closed form
  learning          training drives the relative L2 loss down on a smooth
                    target function
"""
import numpy as np
import pytest

from test_nrc_oracle import random_params, random_raw

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1])
def test_mfma_lane_layout(pkg, gpu_ctx, precision):
    """The operand layout the cache kernels assume for v_mfma_f32_16x16x32:
    small integers (exact in fp16 and bf16), so C must equal A B exactly; a
    K-slot mix-up between lane groups would scramble the products."""
    rng = np.random.default_rng(precision)
    a = rng.integers(-4, 5, (16, 32)).astype(np.float32)
    b = rng.integers(-4, 5, (32, 16)).astype(np.float32)
    c = pkg.vrt.mfma_probe(gpu_ctx, a, b, precision)
    np.testing.assert_array_equal(c, a @ b)


def _layer_slices():
    nw = 8192 + 4 * 4096 + 1024
    out = []
    offs = [(0, 64, 128)] + [(8192 + (l - 1) * 4096, 64, 64) for l in range(1, 5)] + \
           [(8192 + 4 * 4096, 16, 64)]
    for l, (off, o, i) in enumerate(offs):
        out.append((f"W{l}", slice(off, off + o * i)))
        out.append((f"b{l}", slice(nw + 64 * l, nw + 64 * l + o)))
    return out


@pytest.mark.parametrize("precision", [0, 1])
def test_inference_matches_oracle(pkg, oracle, gpu_ctx, precision):
    nrc = pkg.vrt.Nrc(gpu_ctx, precision=precision)
    try:
        params = random_params(11)
        nrc.set_params(params)
        raw = random_raw(5000, 12)
        g = nrc.infer(raw)
        o = oracle.nrc_forward(params, raw, precision)
        scale = np.abs(o).max()
        err = np.abs(g - o).max() / scale
        print("inference max error / scale", err)
        assert err <= (0.01 if precision == 0 else 0.05)
    finally:
        nrc.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_gradients_match_oracle(pkg, oracle, gpu_ctx, precision):
    nrc = pkg.vrt.Nrc(gpu_ctx, precision=precision, max_batch=4096)
    try:
        params = random_params(21, scale=0.8)
        nrc.set_params(params)
        raw = random_raw(3000, 22)  # not a multiple of 64: ragged last chunk
        target = np.random.default_rng(23).uniform(0, 1, (3000, 3)).astype(np.float32)
        g, loss = nrc.gradients(raw, target)
        go, loss_o = oracle.nrc_gradients(params, raw, target, precision)
        assert abs(loss - loss_o) <= 0.02 * abs(loss_o)
        tol = 0.03 if precision == 0 else 0.1
        for name, sl in _layer_slices():
            num = np.linalg.norm(g[sl] - go[sl])
            den = np.linalg.norm(go[sl]) + 1e-12
            print(name, num / den)
            assert num / den <= tol, name
    finally:
        nrc.close()


def test_adam_step_closed_form(pkg, gpu_ctx):
    vrt = pkg.vrt
    lr, b1, b2, eps = 1e-2, 0.9, 0.99, 1e-8
    nrc = vrt.Nrc(gpu_ctx, learning_rate=lr, beta1=b1, beta2=b2, epsilon=eps, max_batch=1024)
    try:
        params = random_params(31)
        nrc.set_params(params)
        raw = random_raw(1024, 32)
        target = np.random.default_rng(33).uniform(0, 1, (1024, 3)).astype(np.float32)
        g, _ = nrc.gradients(raw, target)
        dr, dt = gpu_ctx.upload(raw), gpu_ctx.upload(target)
        nrc.train_step(dr, dt, 1024)
        new = nrc.params()
        gm = g.astype(np.float64) / 1024
        m = (1 - b1) * gm
        v = (1 - b2) * gm * gm
        exp = params - lr * (m / (1 - b1)) / (np.sqrt(v / (1 - b2)) + eps)
        np.testing.assert_allclose(new, exp, rtol=2e-5, atol=2e-7)
        dr.free()
        dt.free()
    finally:
        nrc.close()


def test_adam_step_count_skips_empty_batches(pkg, gpu_ctx):
    """Mini-batches without samples leave Adam untouched, its bias-correction
    step count included: after a cache frame whose camera sees no volume
    (16 empty mini-batches), one training step still follows the first-step
    closed form."""
    import dataclasses
    vrt, wls = pkg.vrt, pkg.workloads
    lr, b1, b2, eps = 1e-2, 0.9, 0.99, 1e-8
    nrc = vrt.Nrc(gpu_ctx, learning_rate=lr, beta1=b1, beta2=b2, epsilon=eps, max_batch=1024)
    try:
        wl = wls.fbm_pathtrace(n=16, width=16, height=8, frames=1, brick=8)
        away = tuple(-x for x in wl.camera["direction"])
        wl = dataclasses.replace(wl, camera=dict(wl.camera, direction=away))
        gs = wls.GpuScene(gpu_ctx, wl)
        opts = gs.nrc_options(train=True, train_samples=256, train_steps=16, query_vertex=1)
        gs.render_nrc(nrc, opts, 0)
        assert nrc.frame_counts()[1] == 0  # no training sample this frame
        params = random_params(35)
        nrc.set_params(params)
        raw = random_raw(1024, 36)
        target = np.random.default_rng(37).uniform(0, 1, (1024, 3)).astype(np.float32)
        g, _ = nrc.gradients(raw, target)
        dr, dt = gpu_ctx.upload(raw), gpu_ctx.upload(target)
        nrc.train_step(dr, dt, 1024)
        new = nrc.params()
        gm = g.astype(np.float64) / 1024
        m = (1 - b1) * gm
        v = (1 - b2) * gm * gm
        exp = params - lr * (m / (1 - b1)) / (np.sqrt(v / (1 - b2)) + eps)
        np.testing.assert_allclose(new, exp, rtol=2e-5, atol=2e-7)
        dr.free()
        dt.free()
    finally:
        nrc.close()


def test_training_reduces_loss(pkg, gpu_ctx):
    vrt = pkg.vrt
    nrc = vrt.Nrc(gpu_ctx, max_batch=16384, seed=3)
    try:
        rng = np.random.default_rng(41)
        raw = rng.uniform(0, 1, (16384 * 4, 6)).astype(np.float32)
        # smooth radiance-like target of position, direction and density
        t = np.stack([0.5 + 0.4 * np.sin(3 * raw[:, 0]) * raw[:, 5],
                      0.3 + 0.2 * np.cos(2 * raw[:, 1] + raw[:, 3]),
                      0.2 + 0.3 * raw[:, 2] * raw[:, 4]], 1).astype(np.float32)
        dr, dt = gpu_ctx.upload(raw), gpu_ctx.upload(t)
        losses = []
        for step in range(200):
            off = (step % 4) * 16384
            sub_r = vrt.DeviceBuffer.__new__(vrt.DeviceBuffer)
            sub_r.ctx, sub_r.ptr, sub_r.nbytes = gpu_ctx, dr.ptr + off * 24, 16384 * 24
            sub_t = vrt.DeviceBuffer.__new__(vrt.DeviceBuffer)
            sub_t.ctx, sub_t.ptr, sub_t.nbytes = gpu_ctx, dt.ptr + off * 12, 16384 * 12
            losses.append(nrc.train_step(sub_r, sub_t, 16384) / 16384)
        print("loss first/last", losses[0], losses[-1])
        assert np.mean(losses[-10:]) < 0.1 * np.mean(losses[:5])
        pred = nrc.infer(raw[:4096])
        rel = np.mean((pred - t[:4096]) ** 2 / (t[:4096] ** 2 + 0.01))
        assert rel < 0.02
    finally:
        nrc.close()
