"""Pins of the CPU oracle against facts that do not come from the oracle.

The reference ships no tests or golden images, so each oracle function is
checked against a published known answer, a closed form, an invariant or a
brute-force enumeration chosen so that a plausible slip (dropped term, wrong
sign or index, transposed operand) fails:

  RNG (common.h:55-102)      TEA round structure vs the published TEA test
                             vector; LCG vs the Numerical Recipes sequence
  TF table (TF1D.cpp:66-111) hand-derived tables for mixed array lengths
  tex3D (A10)                affine fields reproduced exactly; texel centres;
                             clamp addressing
  dda3 (DDA.h)               brute-force slab enumeration of every cell
  majorants (A9)             max over 10^5 random points never exceeds them
  delta tracking (A7)        homogeneous medium: escape probability exp(-sL)
                             and truncated-exponential mean depth
  raycast (A14/A15)          constant medium closed form with the alpha
                             cut-off and the double premultiply
  accumulation (A20)         running mean equals arithmetic mean
  unbiasedness               images with and without majorant grid agree
"""
import ctypes as C
import math

import numpy as np
import pytest

M32 = 0xFFFFFFFF


# ---------------------------------------------------------------------------
# RNG
# ---------------------------------------------------------------------------
def tea(v0, v1, key, rounds):
    s = 0
    for _ in range(rounds):
        s = (s + 0x9E3779B9) & M32
        v0 = (v0 + ((((v1 << 4) & M32) + key[0]) ^ (v1 + s) ^ ((v1 >> 5) + key[1]))) & M32
        v1 = (v1 + ((((v0 << 4) & M32) + key[2]) ^ (v0 + s) ^ ((v0 >> 5) + key[3]))) & M32
    return v0, v1


def test_tea_round_structure_published_vector():
    # TEA (Wheeler & Needham 1994), 32 rounds, zero key and plaintext
    assert tea(0, 0, (0, 0, 0, 0), 32) == (0x41EA3A0A, 0x94BAA940)


def test_lcg_numerical_recipes_sequence(oracle):
    # Numerical Recipes quick-and-dirty generator from state 0
    expect = [1013904223, 1196435762, 3519870697, 2868466484, 1649599747]
    s, got = 0, []
    for _ in range(5):
        s = (1664525 * s + 1013904223) & M32
        got.append(s)
    assert got == expect
    # the oracle emits (state & 0xFFFFFF) / 2^24 of that sequence after seeding
    key = (0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E)
    for seed in [(0, 0), (1, 0), (123456, 7), (2073599, 63)]:
        st = tea(seed[0], seed[1], key, 4)[0]
        assert oracle.lib().orc_tea_seed(*seed) == st
        vals = oracle.lcg_stream(seed[0], seed[1], 16)
        exp = []
        for _ in range(16):
            st = (1664525 * st + 1013904223) & M32
            exp.append((st & 0xFFFFFF) / 16777216.0)
        np.testing.assert_array_equal(vals, np.array(exp, np.float32))


# ---------------------------------------------------------------------------
# Transfer function table
# ---------------------------------------------------------------------------
def test_tf_build_tables(oracle):
    # uniform colour and opacity: one entry
    t = oracle.tf_build(uniform=(0.2, 0.4, 0.6, 0.5))
    np.testing.assert_allclose(t, [[0.2, 0.4, 0.6, 0.5]])
    # 2 RGB colours, 3 opacities: 3 entries, colours lerped at the midpoint,
    # alpha = uniform opacity (1) * opacity array value
    cols = np.array([[1, 0, 0], [0, 0, 1]], np.float32)
    op = np.array([0.0, 0.5, 1.0], np.float32)
    t = oracle.tf_build(cols, op)
    np.testing.assert_allclose(t, [[1, 0, 0, 0.0], [0.5, 0, 0.5, 0.5], [0, 0, 1, 1.0]],
                               atol=1e-7)
    # RGBA colours keep their own alpha and are scaled by the opacity array
    cols4 = np.array([[0, 1, 0, 0.5], [0, 1, 0, 0.5]], np.float32)
    t = oracle.tf_build(cols4, np.array([0.2, 0.2], np.float32))
    np.testing.assert_allclose(t[:, 3], [0.1, 0.1], atol=1e-7)


# ---------------------------------------------------------------------------
# Trilinear fetch + classification
# ---------------------------------------------------------------------------
def _affine_volume(oracle, dims=(9, 7, 5)):
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    vox = (0.25 + 0.5 * x / nx + 0.125 * y / ny - 0.0625 * z / nz).astype(np.float32)
    tf = np.stack([np.linspace(0, 1, 8)] * 4, 1).astype(np.float32)
    return oracle.OracleVolume(vox, (0, 0, 0), (1, 1, 1), tf, (0, 1), 1.0), dims


def test_trilinear_reproduces_affine_field(oracle):
    vol, (nx, ny, nz) = _affine_volume(oracle)
    rng = np.random.default_rng(1)
    # texel coordinate c = p*N - 0.5 must stay inside [0, N-1] to avoid clamping
    c = rng.uniform(0, 1, (5000, 3)) * (np.array([nx, ny, nz]) - 1)
    p = (c + 0.5) / np.array([nx, ny, nz])
    got = vol.sample(p)
    exp = 0.25 + 0.5 * c[:, 0] / nx + 0.125 * c[:, 1] / ny - 0.0625 * c[:, 2] / nz
    assert np.max(np.abs(got - exp)) < 2e-6


def test_trilinear_texel_centres_and_clamp(oracle):
    vol, (nx, ny, nz) = _affine_volume(oracle)
    vox = vol.voxels
    ijk = np.array([[0, 0, 0], [nx - 1, ny - 1, nz - 1], [3, 2, 1], [8, 0, 4]])
    p = (ijk + 0.5) / np.array([nx, ny, nz])
    np.testing.assert_array_equal(vol.sample(p), vox[ijk[:, 2], ijk[:, 1], ijk[:, 0]])
    # far outside the [0,1]^3 box the fetch clamps to the corner voxels
    far = np.array([[-3, -3, -3], [4, 4, 4], [-1, 0.5 / ny, 9]], np.float64)
    np.testing.assert_array_equal(vol.sample(far),
                                  [vox[0, 0, 0], vox[-1, -1, -1], vox[-1, 0, 0]])


def test_classify_normalises_value_range(oracle):
    tf = np.array([[0, 0, 0, 0], [1, 1, 1, 1]], np.float32)
    vol = oracle.OracleVolume(np.zeros((2, 2, 2), np.float32), (0, 0, 0), (1, 1, 1), tf,
                              (2.0, 6.0), 1.0)
    # value 4 -> u = 0.5 -> texel coordinate 0.5 -> halfway between entries
    out = vol.classify([4.0, 2.0, 6.0, 100.0, -100.0])
    np.testing.assert_allclose(out[:, 3], [0.5, 0.0, 1.0, 1.0, 0.0], atol=1e-7)


# ---------------------------------------------------------------------------
# DDA
# ---------------------------------------------------------------------------
def _dda(oracle, ori, d, tmin, tmax, dims):
    cells = np.zeros(4096, np.int32)
    t0 = np.zeros(4096, np.float32)
    t1 = np.zeros(4096, np.float32)
    o = np.asarray(ori, np.float32)
    dd = np.asarray(d, np.float32)
    dm = np.asarray(dims, np.int32)
    lo = np.zeros(3, np.float32)
    hi = np.ones(3, np.float32)
    n = oracle.lib().orc_dda_cells(o.ctypes.data, dd.ctypes.data, tmin, tmax, dm.ctypes.data,
                                   lo.ctypes.data, hi.ctypes.data, 4096, cells.ctypes.data,
                                   t0.ctypes.data, t1.ctypes.data)
    return cells[:n], t0[:n], t1[:n]


def test_dda_matches_bruteforce_enumeration(oracle):
    rng = np.random.default_rng(3)
    dims = np.array([5, 4, 7])
    tested = cells_seen = 0
    for _ in range(300):
        o = rng.uniform(-0.5, 1.5, 3)
        target = rng.uniform(0.05, 0.95, 3)
        d = target - o
        d /= np.linalg.norm(d)
        # clip to the unit box like intersect(Ray, Volume) does before dda3
        inv = 1 / d
        ta, tb = (0 - o) * inv, (1 - o) * inv
        tn, tf = max(np.minimum(ta, tb).max(), 0.0), np.maximum(ta, tb).min()
        if tf <= tn:
            continue
        cells, t0, t1 = _dda(oracle, o, d, tn, tf, dims)
        # brute force: every cell whose slab interval overlaps (tn, tf)
        ref = []
        for cz in range(dims[2]):
            for cy in range(dims[1]):
                for cx in range(dims[0]):
                    lo = np.array([cx, cy, cz]) / dims
                    hi = (np.array([cx, cy, cz]) + 1) / dims
                    a, b = (lo - o) * inv, (hi - o) * inv
                    e, x = max(np.minimum(a, b).max(), tn), min(np.maximum(a, b).min(), tf)
                    if x - e > 1e-5:
                        ref.append((e, (cz * dims[1] + cy) * dims[0] + cx, x))
        ref.sort()
        inside = t1 - t0 > 1e-5
        assert [c for c in cells[inside]] == [c for _, c, _ in ref]
        np.testing.assert_allclose(t0[inside], [e for e, _, _ in ref], atol=2e-5)
        np.testing.assert_allclose(t1[inside], [x for _, _, x in ref], atol=2e-5)
        # cells reported beyond tmax are empty segments at tmax (DDA.h:75-78)
        assert np.all(np.abs(t0[~inside] - t1[~inside]) <= 1e-5)
        tested += 1
        cells_seen += len(ref)
    assert tested > 150 and cells_seen > 5 * tested


# ---------------------------------------------------------------------------
# Majorants
# ---------------------------------------------------------------------------
def test_majorants_bound_every_sample(pkg, oracle):
    vox = pkg.vrt.gen_fbm((40, 36, 33), seed=9)
    tf = pkg.workloads.albedo_tf(0.8, 32)
    tf[:, 3] = np.sin(np.linspace(0, 7, 32)) ** 2  # non-monotone opacity
    vol = oracle.OracleVolume(vox, (0, 0, 0), (1, 1, 1), tf, (0, 1), 1.0)
    G = (5, 4, 3)
    maj = vol.build_majorants(G)
    rng = np.random.default_rng(4)
    p = rng.uniform(0, 1, (100000, 3))
    a = vol.classify(vol.sample(p))[:, 3]
    cell = np.minimum((p * np.array(G)).astype(int), np.array(G) - 1)
    m = maj[cell[:, 2], cell[:, 1], cell[:, 0]]
    assert np.all(a <= m + 1e-6)
    # and the grid is not trivially loose: some sample gets within 5% of it
    assert np.mean(a >= 0.95 * m) > 0.001


# ---------------------------------------------------------------------------
# Delta tracking
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("fp64", [0, 1])
def test_delta_tracking_homogeneous(oracle, fp64):
    vox = np.full((8, 8, 8), 0.5, np.float32)
    tf = np.stack([np.full(4, 0.8)] * 3 + [np.full(4, 0.3)], 1).astype(np.float32)
    ud = 0.25
    vol = oracle.OracleVolume(vox, (-0.5,) * 3, (1 / 7,) * 3, tf, (0, 1), ud)
    vol.attach_majorants((2, 2, 2))
    sigma = 0.3 / ud
    n = 40000
    rng = np.random.default_rng(2)
    esc, depth = 0, []
    st = C.c_uint32(0)
    t = C.c_float()
    d = np.array([0, 0, 1], np.float32)
    for i in range(n):
        o = np.array([rng.uniform(-0.4, 0.4), rng.uniform(-0.4, 0.4), -2.0], np.float32)
        st.value = int(oracle.lib().orc_tea_seed(i, 17))
        hit = oracle.lib().orc_delta_track(C.byref(vol.s), o.ctypes.data, d.ctypes.data, 0.0,
                                           1e30, C.byref(st), C.byref(t), fp64)
        if hit:
            depth.append(t.value - 1.5)
        else:
            esc += 1
    p = math.exp(-sigma)
    assert abs(esc / n - p) <= 5 * math.sqrt(p * (1 - p) / n)
    depth = np.array(depth)
    m = 1 / sigma - math.exp(-sigma) / (1 - math.exp(-sigma))
    assert abs(depth.mean() - m) <= 5 * depth.std() / math.sqrt(depth.size)


# ---------------------------------------------------------------------------
# Raycast closed form
# ---------------------------------------------------------------------------
def test_raycast_constant_medium_closed_form(pkg, oracle):
    a, rgb, ud = 0.05, (0.9, 0.5, 0.1), 0.02
    n = 9
    vox = np.full((n, n, n), 0.5, np.float32)
    tf = np.array([list(rgb) + [a]] * 4, np.float32)
    vol = oracle.OracleVolume(vox, (-0.5,) * 3, (1 / (n - 1),) * 3, tf, (0, 1), ud)
    scene = oracle.OracleScene([vol])
    W = H = 8
    cam = oracle.camera((0.1, 0.2, 3.0), (0, 0, -1), (0, 1, 0), math.radians(10), 1.0)
    bg = (0.1, 0.2, 0.3, 1.0)
    for rate in (2.0, 0.25):
        prm = oracle.params(background=bg, volume_sampling_rate=rate)
        out = np.empty((H, W, 4), np.float32)
        oracle.lib().orc_raycast(C.byref(scene.s), C.byref(cam), C.byref(prm), W, H,
                                 out.ctypes.data, 1)
        dt = np.float32(1 / (n - 1)) * np.float32(1 / rate)
        for y in range(H):
            for x in range(W):
                o = np.zeros(3, np.float32)
                dd = np.zeros(3, np.float32)
                oracle.lib().orc_camera_ray(C.byref(cam), x, y, W, H, o.ctypes.data,
                                            dd.ctypes.data)
                inv = 1 / dd.astype(np.float64)
                ta, tb = (-0.5 - o) * inv, (0.5 - o) * inv
                tn, tf_ = np.minimum(ta, tb).max(), np.maximum(ta, tb).min()
                # number of marching steps with t <= tmax, float32 stepping
                t, k = np.float32(max(tn, 0)), 0
                while t <= np.float32(tf_):
                    k += 1
                    t = np.float32(t + dt)
                # closed form with the alpha >= 0.99 cut-off
                step_t = (1 - a) ** (float(dt) / ud)
                k_cut = math.ceil(math.log(0.01) / math.log(step_t))
                k = min(k, k_cut)
                alpha = 1 - step_t ** k
                color = np.array(rgb) * alpha
                exp = np.append(alpha * color, alpha) + (1 - alpha) * np.array(bg)
                np.testing.assert_allclose(out[y, x], exp, atol=2e-5)


def test_pow_ieee_error_bound(oracle):
    L = oracle.lib()
    L.orc_pow_ieee.restype = C.c_float
    L.orc_pow_ieee.argtypes = [C.c_float, C.c_float]
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.random(20000), 10 ** rng.uniform(-30, 0, 20000)]).astype(np.float32)
    ys = (rng.random(40000) * 4).astype(np.float32)
    got = np.array([L.orc_pow_ieee(float(x), float(y)) for x, y in zip(xs, ys)], np.float64)
    ref = np.power(xs.astype(np.float64), ys.astype(np.float64))
    ok = ref > 1e-37
    z = np.abs(ys * np.log2(xs.astype(np.float64)))
    ulp = np.abs(got - ref) / np.spacing(ref.astype(np.float32)).astype(np.float64)
    assert np.all(ulp[ok] <= 1.5 * (2 + z[ok]))
    assert L.orc_pow_ieee(0.0, 0.5) == 0.0 and L.orc_pow_ieee(1.0, 3.0) == 1.0
    assert L.orc_pow_ieee(0.25, 0.5) == 0.5
    # exponents 0.5 and 2 are the correctly rounded square root and square
    xs = rng.random(5000).astype(np.float32) + np.float32(1e-6)
    half = np.array([L.orc_pow_ieee(float(x), 0.5) for x in xs], np.float32)
    np.testing.assert_array_equal(half, np.sqrt(xs.astype(np.float64)).astype(np.float32))
    two = np.array([L.orc_pow_ieee(float(x), 2.0) for x in xs], np.float32)
    np.testing.assert_array_equal(two, (xs.astype(np.float64) ** 2).astype(np.float32))
    # x near 1 (the opacity correction of small opacities) keeps its digits
    xs = (1.0 - np.geomspace(1e-7, 1e-2, 2000)).astype(np.float32)
    got = np.array([L.orc_pow_ieee(float(x), 0.5) for x in xs], np.float64)
    ref = np.power(xs.astype(np.float64), 0.5)
    assert np.all(np.abs(got - ref) / np.spacing(ref.astype(np.float32)) <= 3.0)


def test_pow_ieee_table(oracle):
    """The log2 table of pow_ieee: subinterval i of z in [0.69921875,
    1.3984375) spans float bits 0x3f330000 + [i, i+1) * 2^17; entry i is
    {float(1/c_i), float(-log2(float(1/c_i)))} with c_i its midpoint, or 1
    for the subinterval holding 1.0."""
    L = oracle.lib()
    tab = np.empty(128, np.float32)
    L.orc_pow_log2_table(tab.ctypes.data)
    tab = tab.reshape(64, 2)
    bits = 0x3f330000 + (np.arange(65, dtype=np.uint32) << 17)
    edges = bits.view(np.float32).astype(np.float64)
    for i in range(64):
        lo, hi = edges[i], edges[i + 1]
        c = 1.0 if lo <= 1.0 < hi else (lo + hi) / 2
        invc = np.float32(1.0 / c)
        assert tab[i, 0] == invc, i
        assert tab[i, 1] == np.float32(-np.log2(np.float64(invc))), i


# ---------------------------------------------------------------------------
# Path tracer invariants
# ---------------------------------------------------------------------------
def _small_pt(pkg, oracle, grid=True, n=24, w=40, h=30):
    wl = pkg.workloads.fbm_pathtrace(n=n, width=w, height=h, frames=1, brick=8)
    if not grid:
        wl.grid_dims = None
    return oracle.OracleWorkload(wl)


def test_accumulation_is_running_mean(pkg, oracle):
    ow = _small_pt(pkg, oracle)
    frames = [ow.pathtrace_frame(a)[0] for a in range(6)]
    acc = ow.pathtrace_accumulate(6)
    np.testing.assert_allclose(acc, np.mean(frames, axis=0), rtol=1e-5, atol=1e-6)


def test_majorant_grid_does_not_bias(pkg, oracle):
    """Delta tracking is unbiased for any majorant >= sigma: the FP64 image mean
    with the macrocell grid equals the one with the single majorant-1 segment."""
    a = _small_pt(pkg, oracle, grid=True).pathtrace_accumulate(128, fp64=True)
    b = _small_pt(pkg, oracle, grid=False).pathtrace_accumulate(128, fp64=True,
                                                                first_accum_id=1000)
    da = a[..., 0].ravel()
    db = b[..., 0].ravel()
    diff = da.mean() - db.mean()
    # per-pixel differences are independent: standard error of their mean
    se = np.std(da - db) / math.sqrt(da.size)
    assert abs(diff) <= 5 * se + 1e-4, (diff, se)


def test_fp32_and_fp64_agree_statistically(pkg, oracle):
    ow = _small_pt(pkg, oracle)
    a = ow.pathtrace_accumulate(64)
    b = ow.pathtrace_accumulate(64, fp64=True)
    err = np.mean((a[..., :3] - b[..., :3]) ** 2 / (b[..., :3] ** 2 + 1e-2))
    assert err <= 1e-3


def test_clip_planes(pkg, oracle):
    """clipRay (renderer/common.h:1469-1482) keeps n.x <= d on every ray.  A
    plane that keeps the whole volume changes neither renderer's image, one
    that keeps none of it leaves the background, and the plane z <= 0 through a
    constant medium starts the march on the plane (closed form)."""
    import dataclasses
    wls = pkg.workloads
    for wl in (wls.blobs_raycast(n=16, width=16, height=12),
               wls.fbm_pathtrace(n=16, width=16, height=12, frames=1, brick=8)):
        def render(planes):
            w = dataclasses.replace(wl, params={**wl.params, "clip_planes": planes})
            ow = oracle.OracleWorkload(w)
            return ow.raycast() if wl.renderer == "raycast" else ow.pathtrace_frame(3)[0]
        base = render(())
        np.testing.assert_array_equal(render([(0, 0, 1, 10.0), (1, 0, 0, 9.0)]), base)
        bg = np.broadcast_to(np.asarray(wl.params["background"], np.float32), base.shape)
        np.testing.assert_array_equal(render([(0, 0, 1, -10.0)]), bg)
        assert not np.array_equal(base, bg)

    a, rgb, ud = 0.05, (0.9, 0.5, 0.1), 0.02
    n = 9
    vox = np.full((n, n, n), 0.5, np.float32)
    tf = np.array([list(rgb) + [a]] * 4, np.float32)
    vol = oracle.OracleVolume(vox, (-0.5,) * 3, (1 / (n - 1),) * 3, tf, (0, 1), ud)
    scene = oracle.OracleScene([vol])
    W = H = 8
    cam = oracle.camera((0.1, 0.2, 3.0), (0, 0, -1), (0, 1, 0), math.radians(10), 1.0)
    bg = (0.1, 0.2, 0.3, 1.0)
    prm = oracle.params(background=bg, volume_sampling_rate=2.0, clip_planes=[(0, 0, 1, 0.0)])
    out = np.empty((H, W, 4), np.float32)
    oracle.lib().orc_raycast(C.byref(scene.s), C.byref(cam), C.byref(prm), W, H,
                             out.ctypes.data, 1)
    dt = np.float32(1 / (n - 1)) * np.float32(1 / 2.0)
    for y in range(H):
        for x in range(W):
            o = np.zeros(3, np.float32)
            dd = np.zeros(3, np.float32)
            oracle.lib().orc_camera_ray(C.byref(cam), x, y, W, H, o.ctypes.data, dd.ctypes.data)
            inv = 1 / dd.astype(np.float64)
            ta, tb = (-0.5 - o) * inv, (0.5 - o) * inv
            tf_ = np.maximum(ta, tb).min()
            # the plane z = 0, reached at t = -o.z / d.z (the ray runs towards -z)
            t, k = np.float32(-o[2]) / dd[2], 0
            while t <= np.float32(tf_):
                k += 1
                t = np.float32(t + dt)
            step_t = (1 - a) ** (float(dt) / ud)
            k = min(k, math.ceil(math.log(0.01) / math.log(step_t)))
            alpha = 1 - step_t ** k
            exp = np.append(alpha * np.array(rgb) * alpha, alpha) + (1 - alpha) * np.array(bg)
            np.testing.assert_allclose(out[y, x], exp, atol=2e-5)
