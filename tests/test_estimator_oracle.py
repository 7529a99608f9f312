"""Absolute pins of the oracle's volume-vertex estimator, from the mathematics
of the light transport rather than from the oracle's own outputs.

Scenes are a homogeneous box [-0.5, 0.5]^3 (constant field, constant TF, so
the extinction is sigma = alpha / unitDistance per world unit) seen head-on
through a narrow camera, rendered with max_bounce 0 so that every pixel
sample is one primary free flight plus the terms of one volume vertex
(Pathtrace_impl.cpp:93-487, 558-600):

* volume NEE of a directional light: albedo * E * Tr(light), with no phase
  factor 1/(4 pi) and no light count (Pathtrace_impl.cpp:330-332), against a
  quadrature of  int sigma e^{-sigma t} e^{-sigma s(t)} dt;
* the ambient term at a volume vertex (Pathtrace_impl.cpp:204-219, 458-466,
  579-583): baseColor * ambientColor * ambientRadiance * vis_ao, with the AO
  ray cosine-distributed around a random normal faceforwarded to the viewer,
  against the closed form for a zero occlusion distance and an independent
  Monte Carlo integration for a finite one;
* the isotropic phase sampler uniform_sample_sphere (Pathtrace_impl.cpp:335);
* Russian roulette (Pathtrace_impl.cpp:594-599) leaves the mean unchanged;
* linear_to_srgb (frame/common.h:43-47) against IEC 61966-2-1 values.
"""
import ctypes as C
import math

import numpy as np
import pytest

BOX_LO, BOX_HI = -0.5, 0.5


def _box_scene(oracle, alpha, unit_distance, albedo, lights):
    vox = np.full((4, 4, 4), 0.5, np.float32)
    tf = np.tile(np.array([[albedo, albedo, albedo, alpha]], np.float32), (8, 1))
    vol = oracle.OracleVolume(vox, (BOX_LO,) * 3, (1.0 / 3.0,) * 3, tf, (0.0, 1.0),
                              unit_distance)
    return oracle.OracleScene([vol], lights), vol


# narrow camera on the box's +z face: the image spans about +-0.05 of the face
CAM_POS = (0.0, 0.0, 3.0)
CAM_FOVY = 2.0 * math.atan(0.05 / 2.5)


def _camera(oracle, w, h):
    return oracle.camera(CAM_POS, (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), CAM_FOVY, w / h)


def _pixel_rays(oracle, cam, w, h):
    """Pixel-centre ray directions (the jitter averages to them to first order)."""
    dirs = np.empty((h, w, 3), np.float64)
    o, d = np.empty(3, np.float32), np.empty(3, np.float32)
    for y in range(h):
        for x in range(w):
            oracle.lib().orc_camera_ray(C.byref(cam), float(x), float(y), w, h, o.ctypes.data,
                                        d.ctypes.data)
            dirs[y, x] = d
    return np.array(CAM_POS, np.float64), dirs


def _box_span(o, d):
    """Entry and exit distance of rays o + t d through the box."""
    with np.errstate(divide="ignore"):
        inv = 1.0 / d
    t1 = (BOX_LO - o) * inv
    t2 = (BOX_HI - o) * inv
    tn = np.max(np.minimum(t1, t2), axis=-1)
    tf = np.min(np.maximum(t1, t2), axis=-1)
    return np.maximum(tn, 0.0), tf


def _exit_distance(p, w):
    """Distance from points p inside the box along unit directions w to its boundary."""
    with np.errstate(divide="ignore", invalid="ignore"):
        tx = np.where(w > 0, (BOX_HI - p) / w, np.where(w < 0, (BOX_LO - p) / w, np.inf))
    return np.min(tx, axis=-1)


def _accumulate(oracle, scene, cam, prm, w, h, frames, first=0):
    out = np.zeros((h, w, 4), np.float64)
    oracle.lib().orc_pathtrace_accumulate(C.byref(scene.s), C.byref(cam), C.byref(prm), w, h,
                                          first, frames, 1, 1, out.ctypes.data, 0)
    return out


def test_volume_nee_directional_quadrature(oracle):
    """Pixel = albedo * E * int_L sigma e^{-sigma t} e^{-sigma s(t)} dt, s(t) the
    distance to the box boundary towards the light; alpha = P(first hit)."""
    alpha, ud, albedo, E = 0.6, 0.4, 0.8, 1.7   # sigma = 1.5 per unit
    sigma = alpha / ud
    light_travel = np.array([0.35, -1.0, 0.55])  # direction the light travels
    L = oracle.light_directional(tuple(light_travel), intensity=E)
    scene, _ = _box_scene(oracle, alpha, ud, albedo, [L])
    w = h = 8
    cam = _camera(oracle, w, h)
    prm = oracle.params(background=(0, 0, 0, 0), ambient_radiance=0.0, ambient_samples=0,
                        max_bounce=0)
    frames = 24000
    img = _accumulate(oracle, scene, cam, prm, w, h, frames)

    o, dirs = _pixel_rays(oracle, cam, w, h)
    t0, t1 = _box_span(o, dirs)
    to_light = -light_travel / np.linalg.norm(light_travel)
    # composite Gauss-Legendre over each pixel's span (the kink where the
    # exit face of the shadow ray changes is resolved by the 128 panels)
    xg, wg = np.polynomial.legendre.leggauss(16)
    expect = np.zeros((h, w))
    for y in range(h):
        for x in range(w):
            a, b = t0[y, x], t1[y, x]
            edges = np.linspace(a, b, 129)
            tot = 0.0
            for lo, hi in zip(edges[:-1], edges[1:]):
                t = 0.5 * (hi - lo) * xg + 0.5 * (hi + lo)
                p = o + t[:, None] * dirs[y, x]
                s = _exit_distance(p, np.broadcast_to(to_light, p.shape))
                f = sigma * np.exp(-sigma * (t - a)) * np.exp(-sigma * s)
                tot += 0.5 * (hi - lo) * np.dot(wg, f)
            expect[y, x] = albedo * E * tot
    p_hit = 1.0 - np.exp(-sigma * (t1 - t0))
    n = w * h * frames
    # each sample is albedo*E*Bernoulli: standard error of the image mean
    se = albedo * E * 0.5 / math.sqrt(n)
    for c in range(3):
        assert abs(img[..., c].mean() - expect.mean()) < 5 * se, (c, img[..., c].mean(),
                                                                  expect.mean())
    assert abs(img[..., 3].mean() - p_hit.mean()) < 5 * 0.5 / math.sqrt(n)
    # an extra 1/(4 pi) or a dropped light transmittance would miss by far
    assert expect.mean() > 0.05 and abs(expect.mean() * (1 - 1 / (4 * math.pi))) > 50 * se


def test_volume_ambient_zero_occlusion_distance(oracle):
    """With occlusionDistance 0 no AO ray is occluded: pixel = albedo *
    ambientColor * ambientRadiance * (1 - e^{-sigma L})."""
    alpha, ud, albedo = 0.5, 0.5, 0.9
    sigma = alpha / ud
    amb_c, amb_r = np.array([1.0, 0.5, 0.25]), 0.7
    scene, _ = _box_scene(oracle, alpha, ud, albedo, [])
    w = h = 8
    cam = _camera(oracle, w, h)
    prm = oracle.params(background=(0, 0, 0, 0), ambient_color=tuple(amb_c),
                        ambient_radiance=amb_r, ambient_samples=1, occlusion_distance=0.0,
                        max_bounce=0)
    frames = 16000
    img = _accumulate(oracle, scene, cam, prm, w, h, frames)
    o, dirs = _pixel_rays(oracle, cam, w, h)
    t0, t1 = _box_span(o, dirs)
    p_hit = (1.0 - np.exp(-sigma * (t1 - t0))).mean()
    n = w * h * frames
    for c in range(3):
        v = albedo * amb_c[c] * amb_r
        assert abs(img[..., c].mean() - v * p_hit) < 5 * v * 0.5 / math.sqrt(n)


def _ao_visibility_mc(rng, o, dirs, sigma, d_occ, n):
    """E[vis_ao] at the collision points of the primary rays (independent Monte
    Carlo): collision depth from the truncated exponential, a normal uniform on
    the hemisphere facing the viewer, the AO direction cosine-distributed about
    it (Malley), visibility e^{-sigma min(d_occ, s)}."""
    h, w, _ = dirs.shape
    pix = rng.integers(0, h * w, n)
    d = dirs.reshape(-1, 3)[pix]
    t0, t1 = _box_span(o, d)
    L = t1 - t0
    u = rng.random(n)
    t = t0 - np.log1p(-u * (1.0 - np.exp(-sigma * L))) / sigma
    p = o + t[:, None] * d
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    view = -d
    nrm = np.where((np.sum(nrm * view, 1) < 0)[:, None], -nrm, nrm)
    # orthonormal frame around the normal by Gram-Schmidt against a fixed axis
    a = np.where((np.abs(nrm[:, 0]) < 0.9)[:, None], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0])
    e1 = a - np.sum(a * nrm, 1, keepdims=True) * nrm
    e1 /= np.linalg.norm(e1, axis=1, keepdims=True)
    e2 = np.cross(nrm, e1)
    r = np.sqrt(rng.random(n))
    phi = 2 * np.pi * rng.random(n)
    z = np.sqrt(np.maximum(0.0, 1.0 - r * r))
    om = (r * np.cos(phi))[:, None] * e1 + (r * np.sin(phi))[:, None] * e2 + z[:, None] * nrm
    s = _exit_distance(p, om)
    return np.exp(-sigma * np.minimum(d_occ, s)), 1.0 - np.exp(-sigma * L)


def test_volume_ambient_occlusion_monte_carlo(oracle):
    """A finite occlusion distance: pixel = albedo * ambient * E[vis_ao], the
    AO ray occluded by a collision within min(occlusionDistance, exit)."""
    alpha, ud, albedo = 0.5, 0.25, 0.9       # sigma = 2
    sigma = alpha / ud
    d_occ = 0.3
    amb_c, amb_r = np.array([1.0, 0.6, 0.3]), 0.8
    scene, _ = _box_scene(oracle, alpha, ud, albedo, [])
    w = h = 8
    cam = _camera(oracle, w, h)
    prm = oracle.params(background=(0, 0, 0, 0), ambient_color=tuple(amb_c),
                        ambient_radiance=amb_r, ambient_samples=1, occlusion_distance=d_occ,
                        max_bounce=0)
    frames = 16000
    img = _accumulate(oracle, scene, cam, prm, w, h, frames)
    o, dirs = _pixel_rays(oracle, cam, w, h)
    vis, phit = _ao_visibility_mc(np.random.default_rng(7), o, dirs, sigma, d_occ, 2_000_000)
    ev = float(np.mean(vis * phit))
    se_mc = float(np.std(vis * phit)) / math.sqrt(len(vis))
    n = w * h * frames
    for c in range(3):
        v = albedo * amb_c[c] * amb_r
        se = v * 0.5 / math.sqrt(n)
        assert abs(img[..., c].mean() - v * ev) < 5 * math.hypot(se, v * se_mc), (
            c, img[..., c].mean(), v * ev)
    # dropping vis_ao (or tracing AO rays through an empty medium) would give
    # albedo * ambient * P(hit) instead
    assert ev < 0.8 * float(np.mean(phit))


def test_uniform_sample_sphere(oracle):
    """The isotropic phase sampler: z = cos(theta) and the azimuth uniform
    (Kolmogorov-Smirnov), unit length, first and second moments of the
    uniform sphere (mean 0, E[w w^T] = I/3)."""
    n = 200000
    out = np.empty((n, 3), np.float32)
    oracle.lib().orc_uniform_sample_sphere(n, 5, out.ctypes.data)
    w = out.astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(w, axis=1), 1.0, atol=2e-6)
    ks_lim = 1.63 / math.sqrt(n) * 1.5
    z = np.sort(w[:, 2])
    assert np.max(np.abs(np.arange(1, n + 1) / n - (z + 1) / 2)) < ks_lim
    phi = np.sort(np.arctan2(w[:, 1], w[:, 0]))
    assert np.max(np.abs(np.arange(1, n + 1) / n - (phi + np.pi) / (2 * np.pi))) < ks_lim
    assert np.all(np.abs(w.mean(0)) < 5 * math.sqrt(1 / 3 / n))
    np.testing.assert_allclose(w.T @ w / n, np.eye(3) / 3, atol=5e-3)


def test_russian_roulette_unbiased(oracle):
    """The image mean is the same with and without Russian roulette (8
    independent batches each; roulette draws from the un-advanced pixel
    stream, so the path streams themselves are the same)."""
    alpha, ud, albedo = 0.6, 0.2, 0.85
    L = oracle.light_directional((0.3, -1.0, -0.4), intensity=1.2)
    scene, _ = _box_scene(oracle, alpha, ud, albedo, [L])
    w = h = 8
    cam = _camera(oracle, w, h)
    prm = oracle.params(background=(0, 0, 0, 0), ambient_radiance=0.3, ambient_samples=1,
                        max_bounce=6)
    batches, frames = 8, 1500

    def batch_means(on):
        oracle.lib().orc_set_roulette(int(on))
        try:
            return np.array([_accumulate(oracle, scene, cam, prm, w, h, frames,
                                         first=b * frames)[..., :3].mean()
                             for b in range(batches)])
        finally:
            oracle.lib().orc_set_roulette(1)

    on, off = batch_means(True), batch_means(False)
    se = math.hypot(on.std(ddof=1), off.std(ddof=1)) / math.sqrt(batches)
    assert abs(on.mean() - off.mean()) < 5 * se + 1e-12, (on.mean(), off.mean(), se)
    # the roulette is active here: it changes the individual batches
    assert not np.array_equal(on, off)


@pytest.mark.parametrize("linear, code", [
    (0.0, 0), (1.0, 254), (0.5, 187), (0.214041, 127), (0.0031308, 10), (0.18, 117),
    (0.001, 3), (2.0, 255), (-1.0, 0)])
def test_srgb_known_values(oracle, linear, code):
    """cvt_uint32_srgb (frame/common.h:31-47): 255 * clamp(linear_to_srgb(x))
    truncated.  IEC 61966-2-1: 0.5 -> 0.7354 (187.5), 0.214041 -> 0.5000,
    the breakpoint 0.0031308 -> 0.04045, 18 % grey -> 0.4614, 12.92 x below
    the breakpoint; alpha passes linearly.  At x >= 1 the curve is 1, but in
    single precision 1.055f * 1 - 0.055f rounds to 1 - 2^-24, so the
    truncated code at exactly 1 is 254 (the clamp acts on the encoded value)."""
    v = np.array([linear, linear, linear, 0.5], np.float32)
    u = oracle.lib().orc_cvt_uint32_srgb(v.ctypes.data)
    assert [(u >> s) & 255 for s in (0, 8, 16)] == [code] * 3
    assert (u >> 24) & 255 == 127


def test_srgb_inverse_brackets(oracle):
    """Every 8-bit code the encoder gives brackets its input under the inverse
    transfer function (sRGB EOTF): EOTF(k/255) <= x < EOTF((k+1)/255)."""
    def eotf(c):
        c = np.asarray(c, np.float64)
        return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)

    xs = np.concatenate([np.linspace(0, 1, 4001), np.geomspace(1e-5, 1, 2000)]).astype(np.float32)
    codes = np.empty(len(xs), np.int64)
    v = np.zeros(4, np.float32)
    for i, x in enumerate(xs):
        v[:] = x
        codes[i] = oracle.lib().orc_cvt_uint32_srgb(v.ctypes.data) & 255
    x64 = xs.astype(np.float64)
    lo, hi = eotf(codes / 255.0), eotf((codes + 1) / 255.0)
    ok = (x64 >= lo * (1 - 1e-5) - 1e-9) & ((x64 < hi * (1 + 1e-5) + 1e-9) | (codes == 255))
    assert ok.all(), xs[~ok][:5]
