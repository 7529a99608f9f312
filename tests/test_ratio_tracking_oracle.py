"""Ratio tracking (an extension for shadow and ambient-occlusion rays) in the
oracle: its transmittance estimate is unbiased -- the mean is exp(-tau) in a
homogeneous box (closed form) and in a heterogeneous field over a majorant
grid (tau by dense quadrature of the pinned field sampler) -- and has less
variance than the binary visibility of delta tracking, which estimates the
same transmittance (DeviceCopyableObjects.h:722-757); the path tracer's image
with ratio tracking has the same expectation as with delta tracking."""
import ctypes as C
import math

import numpy as np
import pytest


def _track(oracle, vol, o, d, tmax, n, fp64, ratio, seed=3):
    out = np.empty(n, np.float64)
    st = C.c_uint32(0)
    t = C.c_float()
    o = np.ascontiguousarray(o, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    for i in range(n):
        st.value = int(oracle.lib().orc_tea_seed(i, seed))
        if ratio:
            out[i] = oracle.lib().orc_ratio_track(C.byref(vol.s), o.ctypes.data, d.ctypes.data,
                                                  0.0, tmax, C.byref(st), fp64)
        else:
            hit = oracle.lib().orc_delta_track(C.byref(vol.s), o.ctypes.data, d.ctypes.data, 0.0,
                                               tmax, C.byref(st), C.byref(t), fp64)
            out[i] = 0.0 if hit else 1.0
    return out


@pytest.mark.parametrize("fp64", [0, 1])
@pytest.mark.parametrize("tmax", [1e30, 2.2])
def test_ratio_tracking_homogeneous(oracle, fp64, tmax):
    """A ray along z through [-0.5, 0.5]^3 from z = -2 (box entry at 1.5):
    E[Tr] = exp(-sigma * (min(tmax, 2.5) - 1.5)).  No majorant grid: the
    majorant is 1 over the box, so each tentative collision keeps 1 - 0.3 of
    the estimate (a tight majorant would make every factor 0, delta
    tracking's binary result)."""
    vox = np.full((8, 8, 8), 0.5, np.float32)
    tf = np.stack([np.full(4, 0.8)] * 3 + [np.full(4, 0.3)], 1).astype(np.float32)
    ud = 0.25
    vol = oracle.OracleVolume(vox, (-0.5,) * 3, (1 / 7,) * 3, tf, (0, 1), ud)
    sigma = 0.3 / ud
    L = min(tmax, 2.5) - 1.5
    n = 20000
    tr = _track(oracle, vol, [0.1, -0.2, -2.0], [0, 0, 1], tmax, n, fp64, True)
    assert np.all((tr >= 0) & (tr <= 1))
    expect = math.exp(-sigma * L)
    assert abs(tr.mean() - expect) <= 5 * tr.std() / math.sqrt(n) + 1e-12, (tr.mean(), expect)
    # per tentative collision the estimate shrinks by 1 - sigma/majorant
    # (0.7 here), so unlike delta tracking it never reaches 0
    assert tr.min() > 0.0 and np.any((tr > 0.0) & (tr < 1.0))


def test_ratio_tracking_heterogeneous_unbiased_lower_variance(oracle):
    """Gaussian blobs over a 4^3 majorant grid, a ray across the volume:
    mean Tr of ratio tracking = exp(-int sigma) = escape rate of delta
    tracking, and its variance is smaller than the Bernoulli variance."""
    rng = np.random.default_rng(11)
    n = 24
    g = np.linspace(0, 1, n)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    vox = np.zeros((n, n, n))
    for _ in range(6):
        c = rng.uniform(0.2, 0.8, 3)
        vox += rng.uniform(0.3, 1.0) * np.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2 +
                                                (z - c[2]) ** 2) / (2 * 0.12 ** 2))
    vox = np.clip(vox, 0, 1).astype(np.float32)
    t = np.linspace(0, 1, 64)
    tf = np.stack([np.full_like(t, 0.7)] * 3 + [t], 1).astype(np.float32)
    ud = 0.3
    vol = oracle.OracleVolume(vox, (-0.5,) * 3, (1 / (n - 1),) * 3, tf, (0, 1), ud)
    vol.attach_majorants((4, 4, 4))
    o = np.array([-2.0, 0.05, -0.1])
    d = np.array([1.0, 0.1, 0.05])
    d /= np.linalg.norm(d)
    # tau by the midpoint rule over the box chord, sampling the pinned field
    # and classification in voxel space
    with np.errstate(divide="ignore"):
        inv = 1 / d
    t0 = np.max(np.minimum((-0.5 - o) * inv, (0.5 - o) * inv))
    t1 = np.min(np.maximum((-0.5 - o) * inv, (0.5 - o) * inv))
    m = 20000
    ts = t0 + (np.arange(m) + 0.5) * (t1 - t0) / m
    pts = (o + ts[:, None] * d) + 0.5  # voxel space [0, 1]^3
    alpha = vol.classify(vol.sample(pts))[:, 3].astype(np.float64)
    tau = alpha.sum() * (t1 - t0) / m / ud
    expect = math.exp(-tau)
    nr = 20000
    tr = _track(oracle, vol, o, d, 1e30, nr, 1, True)
    vis = _track(oracle, vol, o, d, 1e30, nr, 1, False, seed=5)
    assert 0.05 < expect < 0.95
    assert abs(tr.mean() - expect) <= 5 * tr.std() / math.sqrt(nr), (tr.mean(), expect)
    assert abs(vis.mean() - expect) <= 5 * math.sqrt(expect * (1 - expect) / nr)
    assert tr.var() < 0.5 * vis.var()


def test_ratio_tracking_image_expectation(oracle):
    """Path-traced images of a homogeneous box with a directional light and
    ambient occlusion (finite distance): ratio-tracked shadow and AO rays give
    the image of delta tracking in expectation (8 independent batches each),
    with less spread between batches."""
    vox = np.full((4, 4, 4), 0.5, np.float32)
    tf = np.tile(np.array([[0.8, 0.8, 0.8, 0.6]], np.float32), (8, 1))
    vol = oracle.OracleVolume(vox, (-0.5,) * 3, (1 / 3,) * 3, tf, (0, 1), 0.3)
    L = oracle.light_directional((0.3, -1.0, -0.4), intensity=1.2)
    scene = oracle.OracleScene([vol], [L])
    cam = oracle.camera((0.0, 0.0, 3.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0),
                        2 * math.atan(0.3 / 2.5), 1.0)
    w = h = 8
    batches, frames = 8, 600

    def batch_means(ratio):
        prm = oracle.params(background=(0, 0, 0, 0), ambient_radiance=0.4, ambient_samples=1,
                            occlusion_distance=0.4, max_bounce=3, ratio_tracking=ratio)
        out = []
        for b in range(batches):
            img = np.zeros((h, w, 4), np.float64)
            oracle.lib().orc_pathtrace_accumulate(C.byref(scene.s), C.byref(cam), C.byref(prm),
                                                  w, h, b * frames, frames, 1, 1,
                                                  img.ctypes.data, 0)
            out.append(img[..., :3].mean())
        return np.array(out)

    d, r = batch_means(False), batch_means(True)
    se = math.hypot(d.std(ddof=1), r.std(ddof=1)) / math.sqrt(batches)
    assert abs(d.mean() - r.mean()) < 5 * se, (d.mean(), r.mean(), se)
    assert not np.array_equal(d, r)
