"""vrt_pathtrace_noalloc (the path tracer with its work-queue counters in the
library's file-scope device array, allocating nothing) against vrt_pathtrace:
the same images bit for bit, over two calls in a row with the second on a
smaller frame, and a clear error for frames beyond its maximum size."""
import ctypes as C

import numpy as np
import pytest


@pytest.mark.gpu
def test_pathtrace_noalloc_matches_pathtrace(pkg, gpu_ctx):
    vrt, wls = pkg.vrt, pkg.workloads
    for n, w, h in [(48, 200, 136), (32, 72, 40)]:  # the second call is smaller
        print(f"noalloc vs pathtrace: fbm {n}^3 {w}x{h}", flush=True)
        wl = wls.fbm_pathtrace(n=n, width=w, height=h, frames=2, brick=8)
        ref = wls.GpuScene(gpu_ctx, wl)
        dut = wls.GpuScene(gpu_ctx, wl, volume=ref.volume)
        ref.frame.reset()
        dut.frame.reset()
        for accum in range(2):
            ref.render(accum)
            vrt.pathtrace_noalloc(gpu_ctx, dut.scene, dut.camera, dut.params, dut.frame, accum)
            gpu_ctx.sync()
            a, b = ref.frame.color_array(), dut.frame.color_array()
            assert np.array_equal(a, b), (n, accum, float(np.abs(a - b).max()))
            assert np.array_equal(ref.frame.accum_array(), dut.frame.accum_array())
        # the frame has volume pixels: the comparison is not vacuous
        assert float(np.abs(ref.frame.color_array()[..., :3]).max()) > 0.0


@pytest.mark.gpu
def test_pathtrace_noalloc_rejects_oversized_frame(pkg, gpu_ctx):
    vrt, wls = pkg.vrt, pkg.workloads
    wl = wls.fbm_pathtrace(n=16, width=8, height=8, frames=1, brick=8)
    gs = wls.GpuScene(gpu_ctx, wl)
    big = vrt.Frame()
    C.memmove(C.byref(big), C.byref(gs.frame.desc), C.sizeof(vrt.Frame))
    big.width, big.height = 40000, 40000  # 1.6e9 pixels > 2^30
    r = vrt.lib().vrt_pathtrace_noalloc(gpu_ctx.handle, C.byref(gs.scene), C.byref(gs.camera),
                                        C.byref(gs.params), C.byref(big), 0, 0)
    assert r == -1
    assert b"exceeds its maximum" in vrt.lib().vrt_last_error()
