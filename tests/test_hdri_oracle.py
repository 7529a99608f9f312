"""HDRI environment light in the oracle (HDRI.cpp:31-162, sampleCDF.h,
DeviceCopyableObjects.h:2398-2418, Pathtrace_impl.cpp:144-182,
Raycast_impl.cpp:158-173): CDF construction, the sampled cells' distribution
and pdf, the direction <-> map-coordinate mappings through the light's frame,
and closed forms under a constant map."""
import ctypes as C
import dataclasses
import math

import numpy as np
import pytest


def _map(w=8, h=4, seed=3, lo=0.05, hi=3.0):
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, (h, w, 3)).astype(np.float32)


def _sample(oracle, L, n, seed=7):
    d = np.empty((n, 3), np.float32)
    p = np.empty(n, np.float32)
    c = np.empty((n, 3), np.float32)
    oracle.lib().orc_hdri_sample(C.byref(L), n, seed, d.ctypes.data, p.ctypes.data, c.ctypes.data)
    return d, p, c


def _eval(oracle, L, dirs):
    dirs = np.ascontiguousarray(dirs, np.float32)
    n = len(dirs)
    rgb = np.empty((n, 3), np.float32)
    pdf = np.empty(n, np.float32)
    uv = np.empty((n, 2), np.float32)
    oracle.lib().orc_hdri_eval(C.byref(L), n, dirs.ctypes.data, rgb.ctypes.data, pdf.ctypes.data,
                               uv.ctypes.data)
    return rgb, pdf, uv


def test_cdf_tables(oracle):
    """makeCDF: per row the float running sum of max(r, g, b) over the row
    total; the row totals' running sum over their total; a row of zeros
    stays zero."""
    m = _map()
    m[2] = 0.0
    h, w = m.shape[:2]
    L = oracle.light_hdri(m)
    rows = np.empty(h, np.float32)
    cols = np.empty(h * w, np.float32)
    assert oracle.lib().orc_hdri_cdf(C.byref(L), rows.ctypes.data, cols.ctypes.data) == 0
    lum = m.max(axis=2)
    run = np.cumsum(lum, axis=1, dtype=np.float32)
    tot = run[:, -1].copy()
    want = np.where(tot[:, None] != 0, run / np.where(tot == 0, 1, tot)[:, None], run)
    np.testing.assert_array_equal(cols.reshape(h, w), want.astype(np.float32))
    rr = np.cumsum(tot, dtype=np.float32)
    np.testing.assert_array_equal(rows, (rr / rr[-1]).astype(np.float32))
    assert (cols.reshape(h, w)[2] == 0).all()


def test_sampled_cells_follow_luminance(oracle):
    """Cells are drawn in proportion to max(r, g, b); the pdf is the cell's
    probability times w h / 4 pi; the direction is the cell corner's, which
    maps back to the corner's map coordinates (toUV undoes toPolar up to its
    28.67-degree term, 0.00019 in u)."""
    from scipy import stats
    m = _map()
    h, w = m.shape[:2]
    L = oracle.light_hdri(m)
    n = 200000
    d, pdf, rgb = _sample(oracle, L, n)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-5)
    _, _, uv = _eval(oracle, L, d)
    du = 28.670 / 360.0 * math.pi - 0.25
    x = np.rint((uv[:, 0] - du) * w).astype(int) % w
    y = np.rint(np.where(uv[:, 1] <= 0.0011, 0.0, uv[:, 1]) * h).astype(int)
    assert np.abs((uv[:, 0] - du) * w - np.rint((uv[:, 0] - du) * w)).max() < 1e-3
    lum = m.max(axis=2).astype(np.float64).ravel()
    p = lum / lum.sum()
    counts = np.bincount(y * w + x, minlength=w * h).astype(np.float64)
    # row 0 (v = 0) is the pole, where every column has the same direction:
    # one bin for the whole row
    obs = np.concatenate([[counts[:w].sum()], counts[w:]])
    expect = np.concatenate([[p[:w].sum()], p[w:]]) * n
    chi = stats.chisquare(obs, expect * (obs.sum() / expect.sum()))
    assert chi.pvalue > 1e-4, chi
    off = y > 0
    np.testing.assert_allclose(pdf[off], p[(y * w + x)[off]] * w * h / (4 * math.pi), rtol=1e-5)
    pole = np.sort(np.unique(np.round(pdf[~off] * 4 * math.pi / (w * h), 7)))
    np.testing.assert_allclose(pole, np.sort(np.round(p[:w], 7))[-len(pole):], atol=2e-7)


def test_frame_of_up_and_direction(oracle):
    """toWorld = [col1 x direction, -up, col0 x col1] (HDRI.cpp:150-154): with
    up = +x, direction = +z, the map's v = 0 row (toPolar's +y) lies at -up
    and its u = 1/4 column at +z."""
    m = np.ones((8, 16, 3), np.float32)
    L = oracle.light_hdri(m, up=(1, 0, 0), direction=(0, 0, 1))
    _, _, uv = _eval(oracle, L, [[-1, 0, 0], [0, 0, 1], [0, 0, -1]])
    assert uv[0, 1] == pytest.approx(0.001)  # -up: the top row (clamped)
    du = 28.670 / 360.0 * math.pi - 0.25
    # toPolar(u, 1/2) = (-cos(2 pi u - pi/2), 0, sin(2 pi u - pi/2)) in the frame
    for dirv, u in (([0, 0, 1], 0.5), ([0, 0, -1], 0.0)):
        got = _eval(oracle, L, [dirv])[2][0]
        assert got[1] == pytest.approx(0.5, abs=1e-6)
        assert (got[0] - du) % 1.0 == pytest.approx(u, abs=1e-5) or \
            abs((got[0] - du) % 1.0 - 1.0) < 1e-5 and u == 0.0


def test_constant_map_closed_forms(oracle):
    """A constant map: radiance = value * scale everywhere, NEE pdf 1/(4 pi),
    hit-side pdf 1/(2 pi^2 sin theta) with theta the polar angle about -up."""
    c = np.float32(0.7)
    m = np.full((16, 32, 3), c, np.float32)
    L = oracle.light_hdri(m, up=(0, 1, 0), direction=(0, 0, 1), scale=2.5)
    d, pdf, rgb = _sample(oracle, L, 5000)
    np.testing.assert_allclose(rgb, c * np.float32(2.5), rtol=1e-6)
    np.testing.assert_allclose(pdf, 1 / (4 * math.pi), rtol=1e-5)
    rng = np.random.default_rng(1)
    v = rng.normal(size=(2000, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    rgb2, pdf2, _ = _eval(oracle, L, v)
    np.testing.assert_allclose(rgb2, c * np.float32(2.5), rtol=1e-6)
    cos_t = -v[:, 1]  # local y = toLocal d; col1 = -up = -y
    sin_t = np.sqrt(1 - cos_t ** 2)
    ok = sin_t > 1e-3
    np.testing.assert_allclose(pdf2[ok], 1 / (2 * math.pi ** 2 * sin_t[ok]), rtol=2e-4)


def _env_workload(pkg, renderer, radiance, **params):
    wl = pkg.workloads.fbm_pathtrace(n=24, width=40, height=32, frames=1, brick=8)
    if renderer == "raycast":
        wl = dataclasses.replace(wl, renderer="raycast", params=dict(volume_sampling_rate=1.0))
    env = dict(type="hdri", radiance=radiance, up=(0, 1, 0), direction=(0, 0, 1), scale=1.5)
    return dataclasses.replace(wl, lights=[env], params={**wl.params, **params})


@pytest.mark.parametrize("renderer", ["pathtrace", "raycast"])
def test_escaping_rays_see_the_map(pkg, oracle, renderer):
    """Under a constant map every camera ray that leaves the scene untouched
    returns value * scale (path tracer: MIS weight 1 at bounce 0; ray
    caster: the backdrop)."""
    c = np.float32(0.4)
    wl = _env_workload(pkg, renderer, np.full((8, 16, 3), c, np.float32))
    ow = oracle.OracleWorkload(wl)
    img = ow.raycast() if renderer == "raycast" else ow.pathtrace_frame(0)[0]
    corner = img[0, 0]  # the volume does not cover the image corner
    np.testing.assert_allclose(corner, [c * 1.5, c * 1.5, c * 1.5, 1.0], rtol=1e-6)
    assert (img[..., 3] == 1.0).all()
