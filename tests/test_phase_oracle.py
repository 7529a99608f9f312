"""Henyey-Greenstein volume phase function in the oracle (an extension of the
reference's isotropic phase, Pathtrace_impl.cpp:330-339): the sampler against
the HG distribution's closed forms, the density's normalisation, and the
path tracer's NEE weight 4 pi p_HG against the isotropic image."""
import ctypes as C
import math

import numpy as np
import pytest


def _hg_cdf(g, c):
    """P(cos theta <= c) of the HG distribution."""
    return (1 - g * g) / (2 * g) * (1 / np.sqrt(1 + g * g - 2 * g * c) - 1 / (1 + g))


@pytest.mark.parametrize("g", [0.3, -0.6, 0.85])
def test_hg_sampler_distribution(oracle, g):
    d = np.array([0.3, -0.5, 0.8], np.float32)
    d /= np.linalg.norm(d)
    n = 200000
    out = np.empty((n, 3), np.float32)
    oracle.lib().orc_hg_sample(g, d.ctypes.data, n, 9, out.ctypes.data)
    np.testing.assert_allclose(np.linalg.norm(out, axis=1), 1.0, atol=2e-6)
    c = out.astype(np.float64) @ d.astype(np.float64)
    # mean cosine of HG is g; its variance is (1 + 2 g^2)/3 - g^2 ... use the
    # sample's own std for the tolerance
    assert abs(c.mean() - g) <= 5 * c.std() / math.sqrt(n)
    # Kolmogorov-Smirnov distance to the HG CDF
    cs = np.sort(c)
    ks = np.max(np.abs(np.arange(1, n + 1) / n - _hg_cdf(g, cs)))
    assert ks < 1.63 / math.sqrt(n) * 1.5
    # the azimuth around d is uniform: the mean of the perpendicular part vanishes
    perp = out - np.outer(c, d)
    assert np.all(np.abs(perp.mean(0)) < 5 / math.sqrt(n))


@pytest.mark.parametrize("g", [0.0, 0.3, -0.7])
def test_hg_density_normalised(oracle, g):
    """(1/2) int_{-1}^{1} 4 pi p_HG(c) dc = 1 (midpoint rule); 1 everywhere at g = 0."""
    m = 200000
    c = -1 + (np.arange(m) + 0.5) * (2.0 / m)
    vals = np.array([oracle.lib().orc_hg_4pi(g, float(x)) for x in c[::100]])
    w = np.array([oracle.lib().orc_hg_4pi(g, float(x)) for x in c])
    assert abs(0.5 * w.sum() * (2.0 / m) - 1.0) < 1e-6
    if g == 0.0:
        assert np.all(vals == 1.0)


def test_nee_carries_phase_weight(pkg, oracle):
    """First-vertex NEE only (max_bounce 0, no ambient): with the same random
    streams the g = 0.3 image is the isotropic one times 4 pi p_HG(d . L)."""
    import dataclasses
    wl = pkg.workloads.fbm_pathtrace(n=24, width=24, height=18, frames=1, brick=8)
    prm = dict(wl.params, max_bounce=0, ambient_samples=0, ambient_radiance=0.0)
    iso = oracle.OracleWorkload(dataclasses.replace(wl, params=prm)).pathtrace_frame(0)[0]
    g = 0.3
    ani = oracle.OracleWorkload(dataclasses.replace(wl, params=dict(prm, phase_g=g)))
    img = ani.pathtrace_frame(0)[0]
    L = -np.array(wl.lights[0]["direction"], np.float64)
    L /= np.linalg.norm(L)
    cam = ani.cam
    o, dv = np.zeros(3, np.float32), np.zeros(3, np.float32)
    ratio_ok, lit = 0, 0
    for y in range(wl.height):
        for x in range(wl.width):
            if iso[y, x, 0] <= 0:
                continue
            lit += 1
            # the camera ray of this sample: the pixel jitter is the first two
            # numbers of the pixel's stream (Pathtrace_impl.cpp:510-520)
            j = oracle.lcg_stream(x + wl.width * y, 0, 2)
            oracle.lib().orc_camera_ray(C.byref(cam), x + j[0] - 0.5, y + j[1] - 0.5, wl.width,
                                        wl.height, o.ctypes.data, dv.ctypes.data)
            d = dv.astype(np.float64) / np.linalg.norm(dv)
            w = oracle.lib().orc_hg_4pi(g, float(d @ L))
            ratio_ok += abs(img[y, x, 0] / iso[y, x, 0] - w) < 1e-4 * w
    assert lit > 50 and ratio_ok == lit
