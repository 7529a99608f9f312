"""Benchmark: plain 'pathtrace' renderer, Msamples/s at 1080p on one MI355X.

Workload (the single-GPU benchmark configuration): structuredRegular 512^3
fp32 fBm (6 octaves, lacunarity 2, gain 0.5, seed 1) remapped to [0,1],
density scale 50 (unitDistance 1/50, TF alpha = value, albedo 0.8), delta
tracking over a 32^3 majorant grid of 16^3-voxel macrocells, one directional
light plus the constant ambient term, up to 8 bounces, 1920x1080, 1 spp per
frame.  A step is one frame (one launch of the path-tracing kernel over the
image); the timed region holds exactly --steps frames with all inputs
resident in HBM.

Also reported on the same JSON line:
  roofline      volume fetch of the path-tracing kernel: algorithmic bytes
                (32 B per trilinear lookup, counted by the kernel itself)
                per launch / average launch time from HIP events on the
                kernel's stream, against 8 TB/s HBM peak; 'traffic' is the
                HBM bytes per launch measured with rocprofv3 counters when a
                summary for this workload exists under profiles/
  cpu_baseline  the CPU restatement of the same renderer (oracle/, FP32,
                one thread per core over 64x64 tiles like the reference CPU
                device) timed on a bounded band of rows of the same frame
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_LOOKUP = 32
# Henyey-Greenstein anisotropy of the fBm path-tracing workload (and of the
# cloud whose time to relMSE is measured), as the workload is specified
PHASE_G = 0.3
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_pathtrace_fbm512_1080p.json")


def _package():
    import __graft_entry__ as ge
    return ge._package()


def cpu_baseline(wl, target_seconds=12.0):
    """Time the oracle (FP32, all host cores) on whole frames of the same
    workload until about target_seconds have passed (at least one frame)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    ow = O.OracleWorkload(wl)
    threads = os.cpu_count() or 1
    frames = 0
    t0 = time.perf_counter()
    while True:
        ow.pathtrace_frame(frames, threads=threads)
        frames += 1
        dt = time.perf_counter() - t0
        if dt >= target_seconds:
            break
    samples = frames * wl.width * wl.height
    return {
        "value": samples / dt / 1e6,
        "unit": "Msamples/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{frames} frames of {wl.width}x{wl.height} x 1 spp ({dt:.1f} s) of the same "
                  f"workload, FP32 CPU restatement, {threads} threads over 64x64 tiles",
    }


PRECISION_NOTE = {"fp16": "fp16 weights/activations, fp32 accumulation, fp32 Adam",
                  "bf16": "bf16 weights/activations, fp32 accumulation, fp32 Adam"}


def reference_scene(pkg, ctx, wl, volume=None, ref_frames=2048):
    """The plain GPU path tracer accumulated over ref_frames spp in FP64: the
    reference every time-to-relMSE figure is measured against."""
    ref = pkg.workloads.GpuScene(ctx, wl, accum64=True, volume=volume)
    ref.frame.reset()
    for a in range(ref_frames):
        ref.render(a)
    ctx.sync()
    return ref


def nrc_vs_plain(pkg, ctx, wl, volume=None, ref_frames=2048, cap=1024, warmup=4, train_steps=8,
                 target=1e-2, oracle_stride=24, oracle_spp=4096, timed_frames=32,
                 precisions=("fp16",), query_vertex=2, train_samples=65536, ref=None,
                 train_extra=0, seeds=3):
    """Wall time to relMSE <= target of the plain path tracer and of the
    radiance-cache path tracer (its training warm-up included) for each cache
    precision, all measured against an FP64-accumulated plain reference of
    ref_frames spp (`ref`, or made here); frame times on HIP events, the
    plain frames on a frame without the FP64 buffer, as the cache frames are;
    the frames each estimator needs are averaged over `seeds` independent
    runs (pixel streams and cache initialisation), since one run lands a
    frame early or late; the reference's own relMSE against the CPU FP64
    oracle is measured on every oracle_stride-th pixel."""
    vrt, wls = pkg.vrt, pkg.workloads
    if ref is None:
        ref = reference_scene(pkg, ctx, wl, volume, ref_frames)
    volume = ref.volume
    timing = wls.GpuScene(ctx, wl, volume=volume)
    timing.frame.reset()
    for a in range(4):
        timing.render(a)
    ctx.sync()
    ctx.timer_start()
    for a in range(timed_frames):
        timing.render(4 + a)
    plain_ms = ctx.timer_stop() / timed_frames

    # plain: pixel streams independent of the reference's (rng_offset)
    plain_runs, plain_reached = [], True
    for sd in range(seeds):
        gs = wls.GpuScene(ctx, wl, volume=volume, params={"rng_offset": (1 + 4 * sd) << 20})
        gs.frame.reset()
        n, e = 0, float("inf")
        while e > target and n < cap:
            gs.render(n)
            n += 1
            e = vrt.relmse(ctx, gs.frame, ref.frame, ref_frames)
        plain_runs.append(n)
        plain_reached = plain_reached and e <= target
    frames_plain = round(sum(plain_runs) / seeds, 2)

    # radiance cache: train from scratch for `warmup` frames, then accumulate
    # until the target; the run is then repeated from scratch for that many
    # frames (same seeds; the training paths claim their slots by atomics, so
    # the mini-batch order and the cache can differ slightly) without the
    # relMSE checks and timed on HIP events
    scenes = [(wls.GpuScene(ctx, wl, volume=volume, params={"rng_offset": (2 + 4 * sd) << 20}),
               wls.GpuScene(ctx, wl, volume=volume, params={"rng_offset": (3 + 4 * sd) << 20}))
              for sd in range(seeds)]
    opts = scenes[0][0].nrc_options(train=True, train_samples=train_samples,
                                    train_steps=train_steps, query_vertex=query_vertex,
                                    train_extra_vertices=train_extra)

    def nrc_run(precision, frames=None, sd=0):
        warm, gn = scenes[sd]
        nrc = vrt.Nrc(ctx, precision=precision, max_batch=16384, seed=1 + sd)
        ctx.sync()
        ctx.timer_start()
        for a in range(warmup):
            warm.render_nrc(nrc, opts, a)
        gn.frame.reset()
        n, e = 0, float("inf")
        while (n < frames) if frames is not None else (e > target and n < cap):
            gn.render_nrc(nrc, opts, n)
            n += 1
            if frames is None:
                e = vrt.relmse(ctx, gn.frame, ref.frame, ref_frames)
        ms = ctx.timer_stop()
        counts = nrc.frame_counts()
        nrc.close()
        return n, e, ms, counts

    t_plain = frames_plain * plain_ms
    out = {
        "workload": wl.name,
        "target_relmse": target,
        "relmse_eps": 1e-2,
        "reference": f"plain GPU path tracer, {ref_frames} spp, FP64 accumulation",
        "plain": {"frames": frames_plain, "frames_per_seed": plain_runs,
                  "ms_per_frame": round(plain_ms, 4),
                  "ms_to_target": round(t_plain, 2) if plain_reached else None,
                  "relmse": e},
    }
    for i, prec in enumerate(precisions):
        code = {"fp16": vrt.NRC_FP16, "bf16": vrt.NRC_BF16}[prec]
        runs = [nrc_run(code, sd=sd) for sd in range(seeds)]
        nrc_runs = [r[0] for r in runs]
        frames_nrc = round(sum(nrc_runs) / seeds, 2)
        e_nrc, counts = runs[0][1], runs[0][3]
        nrc_reached = all(r[1] <= target for r in runs)
        # frame time: the first seed's run repeated without relMSE checks
        _, _, t_run, _ = nrc_run(code, runs[0][0])
        nrc_ms = t_run / (warmup + runs[0][0])
        t_nrc = (warmup + frames_nrc) * nrc_ms
        key = "nrc" if i == 0 else f"nrc_{prec}"
        out[key] = {"query_vertex": query_vertex, "train_samples_per_frame": train_samples,
                    "train_extra_vertices": train_extra, "warmup_frames": warmup, "frames": frames_nrc,
                    "frames_per_seed": nrc_runs,
                    "ms_per_frame": round(nrc_ms, 4),
                    "ms_to_target": round(t_nrc, 2) if nrc_reached else None, "relmse": e_nrc,
                    "queries_last_frame": counts[0],
                    "training_samples_last_frame": min(counts[1], train_samples),
                    "train_steps_per_frame": train_steps,
                    "precision": PRECISION_NOTE[prec]}
        out["speedup" if i == 0 else f"speedup_{prec}"] = (
            round(t_plain / t_nrc, 3) if plain_reached and nrc_reached else None)
    # the reference against the CPU FP64 oracle on a pixel subset
    if oracle_stride:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        import numpy as np
        t0 = time.perf_counter()
        o = O.OracleWorkload(wl).pathtrace_accumulate(oracle_spp, first_accum_id=50000, fp64=True,
                                                      stride=oracle_stride,
                                                      threads=os.cpu_count() or 1)
        g = (ref.frame.accum64_array() / ref_frames)[::oracle_stride, ::oracle_stride]
        o3, g3 = o[..., :3], g[..., :3]
        out["reference_vs_cpu_fp64_oracle"] = {
            "relmse": float(np.mean((g3 - o3) ** 2 / (o3 ** 2 + 1e-2))),
            "pixels": int(o.shape[0] * o.shape[1]), "oracle_spp": oracle_spp,
            "seconds": round(time.perf_counter() - t0, 1)}
    return out


def mixed_frame_time(pkg, ctx, width, height, frames=32, warmup=4, builder="sah"):
    """Plain path-tracer frame time on the mixed surface + volume scene:
    1,310,720-triangle icosphere, 256^3 blob volume, quad light, <= 6
    bounces, 1 spp per frame; the mesh BVH built on the host by binned SAH
    ("sah") or on the GPU as a linear BVH ("lbvh")."""
    wls = pkg.workloads
    t0 = time.perf_counter()
    wl = wls.mixed_pathtrace(level=8, vol_n=256, width=width, height=height, frames=frames)
    gs = wls.GpuScene(ctx, wl, bvh_builder=builder)
    setup = time.perf_counter() - t0
    info = gs.meshes[0][0].info
    gs.frame.reset()
    for a in range(warmup):
        gs.render(a)
    ctx.sync()
    ctx.timer_start()
    for a in range(frames):
        gs.render(warmup + a)
    ms = ctx.timer_stop() / frames
    return {"workload": wl.name, "bvh_builder": builder, "ms_per_frame": round(ms, 4),
            "Msamples_per_s": round(width * height / (ms / 1e3) / 1e6, 2),
            "triangles": int(info.num_triangles), "bvh_nodes": int(info.num_nodes),
            "bvh_depth": int(info.depth), "bvh_build_ms": round(info.build_ms, 1),
            "setup_s": round(setup, 1)}


def raycast_frame_time(pkg, ctx, n=128, width=512, height=512, frames=32, warmup=4, cpu=True,
                       note="the 8 MiB volume lives in L2/MALL"):
    """The ray caster on an n^3 blob volume (n = 128 at 512x512 is the
    reference's CPU-runnable configuration): GPU frame time on HIP events,
    trilinear lookups per frame (exact: the march is deterministic) and the
    algorithmic fetch rate; with `cpu`, the CPU restatement timed on this
    host's cores for the same frame."""
    wls = pkg.workloads
    wl = wls.blobs_raycast(n=n, width=width, height=height)
    gs = wls.GpuScene(ctx, wl)
    for a in range(warmup):
        gs.render(a)
    ctx.sync()
    gs.frame.counters.zero()
    ctx.sync()
    ctx.timer_start()
    for a in range(frames):
        gs.render(a)
    ms = ctx.timer_stop() / frames
    lookups = int(gs.frame.counter_values()[0]) / frames
    out = {"workload": wl.name, "ms_per_frame": round(ms, 4),
           "Msamples_per_s": round(wl.width * wl.height / (ms / 1e3) / 1e6, 1),
           "lookups_per_frame": int(lookups),
           "achieved_GBps": round(lookups * BYTES_PER_LOOKUP / (ms / 1e3) / 1e9, 1),
           "frac_of_hbm_peak": round(lookups * BYTES_PER_LOOKUP / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
           "note": note}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        threads = os.cpu_count() or 1
        ow = O.OracleWorkload(wl)
        ow.raycast(threads=threads)
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            ow.raycast(threads=threads)
        cms = (time.perf_counter() - t0) / reps * 1e3
        out["cpu_restatement"] = {"ms_per_frame": round(cms, 1), "threads": threads,
                                  "Msamples_per_s": round(wl.width * wl.height / (cms / 1e3) / 1e6,
                                                          2)}
    return out


def calibrate_copy(vrt, mib, reps):
    """Copy kernel of known size: each launch reads and writes mib MiB."""
    ctx = vrt.Context(0)
    n = mib << 20
    a, b = ctx.buffer(n), ctx.buffer(n)
    a.zero()
    for _ in range(2):
        ctx.copy(b, a, n)
    ctx.timer_start()
    for _ in range(reps):
        ctx.copy(b, a, n)
    ms = ctx.timer_stop()
    print(json.dumps({"kernel": "k_copy", "bytes_read_per_launch": n, "bytes_written_per_launch": n,
                      "launches": reps, "ms_per_launch": ms / reps,
                      "GBps_read_plus_write": 2 * n * reps / (ms / 1e3) / 1e9}), flush=True)
    ctx.close()


# radiance-cache network, unpadded multiply-adds per sample: 121x64 + 4 x 64x64
# + 64x3 forward; the backward adds W^T dZ through layers 5..1 (64x3 + 4 x
# 64x64) and dW of every layer (the forward count again)
NRC_FWD_MACS = 121 * 64 + 4 * 64 * 64 + 64 * 3
NRC_BWD_MACS = (64 * 3 + 4 * 64 * 64) + NRC_FWD_MACS
MFMA_PEAK_TFLOPS = {"fp16": 2500.0, "bf16": 2500.0}  # MI355X dense


def traversal_rates(pkg, ctx, n=1 << 21, reps=5, builder="sah"):
    """BVH traversal alone on the mixed scene's 1,310,720-triangle icosphere:
    n incoherent rays from points inside the sphere (uniform in a ball of
    radius 0.3, uniform directions), closest hit; and n rays from the same
    points towards the quad light's area, any hit and closest hit, through
    vrt_intersect_surfaces."""
    import ctypes as C
    import numpy as np
    vrt = pkg.vrt
    v, i = vrt.gen_icosphere(level=8, radius=0.5, amp=0.08, freq=3.0, seed=2)
    mesh = vrt.DeviceMesh(ctx, v, i, builder=builder)
    scene = vrt.make_scene([], [], [(mesh, (0.7, 0.7, 0.7))])
    rng = np.random.default_rng(1)
    o = rng.normal(size=(n, 3))
    o *= 0.3 * rng.uniform(0, 1, (n, 1)) ** (1 / 3) / np.linalg.norm(o, axis=1, keepdims=True)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3], rays[:, 3:6], rays[:, 6], rays[:, 7] = o, d, 1e-4, 1e30
    tgt = np.array([0.0, 0.36, 0.0]) + rng.uniform(-0.08, 0.08, (n, 3)) * np.array([1, 0, 1])
    sd = tgt - o
    dist = np.linalg.norm(sd, axis=1)
    srays = rays.copy()
    srays[:, 3:6], srays[:, 7] = sd / dist[:, None], dist * 0.999
    dr, ds = ctx.upload(rays), ctx.upload(srays)
    dt, dp = ctx.buffer(4 * n), ctx.buffer(4 * n)
    lib = vrt.lib()
    out = {"triangles": len(i), "bvh_builder": builder, "rays": n,
           "bvh_nodes": int(mesh.info.num_nodes)}
    for name, buf, anyh in (("closest", dr, 0), ("light_any", ds, 1), ("light_closest", ds, 0)):
        def run():
            vrt.check(lib.vrt_intersect_surfaces(ctx.handle, C.byref(scene), n, buf.ptr, anyh,
                                                 dt.ptr, dp.ptr, None))
        run()
        ctx.sync()
        ctx.timer_start()
        for _ in range(reps):
            run()
        ms = ctx.timer_stop() / reps
        hit = float((dp.download(np.int32, (n,)) >= 0).mean())
        out[name] = {"ms": round(ms, 4), "Mrays_per_s": round(n / ms / 1e3, 1),
                     "hit_fraction": round(hit, 4)}
    for b in (dr, ds, dt, dp):
        b.free()
    mesh.close()
    return out


def mfma_rates(pkg, ctx, infer_n=1 << 21, train_n=65536, reps=20):
    """The cache's MFMA kernels alone, on device-resident random inputs:
    inference over infer_n samples (a 1080p frame's worth of queries) and the
    gradient pass (forward + backward + reduction) over train_n samples,
    timed on HIP events; algorithmic TFLOP/s from the unpadded multiply-adds
    and the fraction of the MI355X dense fp16/bf16 peak."""
    import numpy as np
    vrt = pkg.vrt
    rng = np.random.default_rng(5)
    raw = ctx.upload(rng.random((infer_n, 6), dtype=np.float32))
    out = ctx.buffer(infer_n * 12)
    tgt = ctx.upload(rng.random((train_n, 3), dtype=np.float32))
    lib = vrt.lib()
    res = {}
    for prec in ("fp16", "bf16"):
        code = {"fp16": vrt.NRC_FP16, "bf16": vrt.NRC_BF16}[prec]
        nrc = vrt.Nrc(ctx, precision=code, max_batch=train_n, seed=3)
        for _ in range(3):
            vrt.check(lib.vrt_nrc_infer(nrc.handle, infer_n, raw.ptr, out.ptr))
            vrt.check(lib.vrt_nrc_gradients(nrc.handle, train_n, raw.ptr, tgt.ptr, None, None))
        ctx.sync()
        ctx.timer_start()
        for _ in range(reps):
            vrt.check(lib.vrt_nrc_infer(nrc.handle, infer_n, raw.ptr, out.ptr))
        ms_inf = ctx.timer_stop() / reps
        ctx.timer_start()
        for _ in range(reps):
            vrt.check(lib.vrt_nrc_gradients(nrc.handle, train_n, raw.ptr, tgt.ptr, None, None))
        ms_grad = ctx.timer_stop() / reps
        nrc.close()
        t_inf = infer_n * NRC_FWD_MACS * 2 / (ms_inf / 1e3) / 1e12
        t_grad = train_n * (NRC_FWD_MACS + NRC_BWD_MACS) * 2 / (ms_grad / 1e3) / 1e12
        res[prec] = {
            "infer": {"samples": infer_n, "ms": round(ms_inf, 4), "TFLOPs": round(t_inf, 1),
                      "frac_of_peak": round(t_inf / MFMA_PEAK_TFLOPS[prec], 4)},
            "train_gradients": {"samples": train_n, "ms": round(ms_grad, 4),
                                "TFLOPs": round(t_grad, 1),
                                "frac_of_peak": round(t_grad / MFMA_PEAK_TFLOPS[prec], 4)}}
    for b in (raw, out, tgt):
        b.free()
    res["peak_TFLOPs"] = MFMA_PEAK_TFLOPS["fp16"]
    res["flops_note"] = (f"unpadded: {2 * NRC_FWD_MACS} FLOP/sample forward, "
                         f"{2 * (NRC_FWD_MACS + NRC_BWD_MACS)} FLOP/sample forward+backward")
    return res


KERNEL_WORKLOADS = ("pathtrace_fbm512", "raycast_blobs512", "raycast_blobs128",
                    "pathtrace_cloud1024", "nrc_cloud1024", "pathtrace_tets2m", "nrc_tets2m",
                    "nrc_mfma", "pathtrace_mixed", "bvh_mixed")


def kernel_only(pkg, ctx, name, steps, warmup, precision="fp16", query_vertex=2,
                train_samples=65536, train_steps=8, train_extra=0, pass_schedule=0):
    """One workload's frames alone (profiling runs): `warmup` untimed frames,
    then `steps` frames timed on HIP events; prints one JSON line.
    `pass_schedule`: the path tracer's vrt_params.pass_schedule (0: per
    scene)."""
    vrt, wls = pkg.vrt, pkg.workloads
    if name == "nrc_mfma":
        print(json.dumps({"kernel_workload": name, **mfma_rates(pkg, ctx, reps=steps)}),
              flush=True)
        return
    if name == "bvh_mixed":
        print(json.dumps({"kernel_workload": name, **traversal_rates(pkg, ctx, reps=steps)}),
              flush=True)
        return
    if name == "pathtrace_fbm512":
        wl = wls.fbm_pathtrace(n=512, phase_g=PHASE_G)
    elif name == "raycast_blobs512":
        wl = wls.blobs_raycast(n=512, width=1920, height=1080)
    elif name == "raycast_blobs128":
        wl = wls.blobs_raycast(n=128)
    elif name in ("pathtrace_cloud1024", "nrc_cloud1024"):
        wl = wls.cloud_nanovdb_pathtrace(n=1024, phase_g=PHASE_G)
    elif name == "pathtrace_mixed":
        wl = wls.mixed_pathtrace(level=8, vol_n=256)
    else:
        wl = wls.tets_pathtrace()
    gs = wls.GpuScene(ctx, wl, params={"pass_schedule": pass_schedule})
    nrc = opts = None
    if name.startswith("nrc_"):
        nrc = vrt.Nrc(ctx, precision={"fp16": vrt.NRC_FP16, "bf16": vrt.NRC_BF16}[precision],
                      max_batch=16384, seed=1)
        opts = gs.nrc_options(train=True, train_samples=train_samples, train_steps=train_steps,
                              query_vertex=query_vertex, train_extra_vertices=train_extra)

    def frame(a):
        if nrc is not None:
            gs.render_nrc(nrc, opts, a)
        else:
            gs.render(a)

    gs.frame.reset()
    for a in range(warmup):
        frame(a)
    ctx.sync()
    gs.frame.counters.zero()
    ctx.sync()
    ctx.timer_start()
    for a in range(steps):
        frame(warmup + a)
    ms = ctx.timer_stop() / steps
    lookups = int(gs.frame.counter_values()[0]) / steps
    out = {"kernel_workload": name, "workload": wl.name, "pass_schedule": pass_schedule,
           "frames": steps,
           "ms_per_frame": round(ms, 4), "lookups_per_frame": lookups,
           "algorithmic_GBps": round(lookups * BYTES_PER_LOOKUP / (ms / 1e3) / 1e9, 1)}
    if nrc is not None:
        out.update(precision=precision, query_vertex=query_vertex, train_samples=train_samples,
                   train_steps=train_steps, train_extra_vertices=train_extra,
                   counts=list(nrc.frame_counts()))
        nrc.close()
    print(json.dumps(out), flush=True)


# what bench.py accepts of a committed traffic profile: it must have been
# taken on a kernel of about the same time and lookups per launch
TRAFFIC_TOLERANCE = 0.10


def load_traffic(ms_per_launch, lookups_per_launch):
    """HBM bytes per launch from the committed rocprofv3 summary of this
    workload, or (None, reason) when it is missing or was taken on a kernel
    whose time or lookup count differs from this run's by more than 10 %."""
    if not os.path.exists(TRAFFIC_FILE):
        return None, "no rocprofv3 traffic summary for this workload"
    with open(TRAFFIC_FILE) as f:
        t = json.load(f)
    pms, plk = t.get("ms_per_launch"), t.get("lookups_per_launch")
    if not pms or not plk:
        return None, "the traffic summary records no kernel time / lookups to check it against"
    if (abs(pms - ms_per_launch) > TRAFFIC_TOLERANCE * ms_per_launch or
            abs(plk - lookups_per_launch) > TRAFFIC_TOLERANCE * lookups_per_launch):
        return None, (f"stale: profiled at {pms:.4f} ms and {plk:.4g} lookups per launch, "
                      f"this run {ms_per_launch:.4f} ms and {lookups_per_launch:.4g}")
    return t.get("hbm_bytes_per_launch"), f"rocprofv3 FETCH_SIZE + WRITE_SIZE, {t.get('note', '')}"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-nrc", action="store_true",
                    help="skip the time-to-relMSE comparisons with the radiance cache")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--pass-schedule", type=int, default=0, choices=(0, 1, 2, 3),
                    help="path-tracer pass schedule of the --kernel runs (0: per scene)")
    ap.add_argument("--nrc-query-vertex", type=int, default=2, choices=(1, 2),
                    help="volume vertex whose cache query ends a path (2: the configuration)")
    ap.add_argument("--nrc-train-extra", type=int, default=0,
                    help="self-training: training paths end in a cache query this many bounces "
                         "after the query vertex (0: traced to full length)")
    ap.add_argument("--ttr-size", type=int, default=1024,
                    help="index-space edge of the nanovdb cloud of the time-to-relMSE run")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the frame time of the mixed triangle + volume scene")
    ap.add_argument("--no-c4", action="store_true",
                    help="skip time-to-relMSE on the 2M-tetrahedron 4K scene (plain vs radiance "
                         "cache in fp16 and bf16)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run the plain copy kernel over MIB MiB (counter calibration)")
    ap.add_argument("--kernel", choices=KERNEL_WORKLOADS, default=None,
                    help="only run this workload's frames (profiling runs)")
    ap.add_argument("--precision", choices=("fp16", "bf16"), default="fp16",
                    help="cache precision of the --kernel nrc_* runs")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the per-pixel path is sharded by image "
              "tiles across GPUs by design, but multi-GPU runs are not part of this build",
              file=sys.stderr)
        sys.exit(2)

    pkg = _package()
    vrt, wls = pkg.vrt, pkg.workloads
    if args.calibrate_copy:
        calibrate_copy(vrt, args.calibrate_copy, args.steps)
        return
    if args.kernel:
        ctx = vrt.Context(0)
        kernel_only(pkg, ctx, args.kernel, args.steps, args.warmup, args.precision,
                    args.nrc_query_vertex, train_extra=args.nrc_train_extra,
                    pass_schedule=args.pass_schedule)
        ctx.close()
        return
    t0 = time.perf_counter()
    wl = wls.fbm_pathtrace(n=args.size, width=args.width, height=args.height,
                           frames=args.steps, phase_g=PHASE_G)
    print(f"# generated {wl.name} in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
          flush=True)

    ctx = vrt.Context(0)
    gs = wls.GpuScene(ctx, wl)
    gs.frame.reset()
    for a in range(args.warmup):
        gs.render(a)
    ctx.sync()
    gs.frame.counters.zero()
    ctx.sync()

    t_wall = time.perf_counter()
    ctx.timer_start()
    for k in range(args.steps):
        gs.render(args.warmup + k)
    ms = ctx.timer_stop()
    ctx.sync()
    wall = time.perf_counter() - t_wall
    cv = gs.frame.counter_values()
    lookups, hit_samples = int(cv[0]), int(cv[1])

    samples = wl.width * wl.height * args.steps
    value = samples / (ms / 1e3) / 1e6
    ms_per_step = ms / args.steps
    lookups_per_launch = lookups / args.steps
    bytes_per_launch = lookups_per_launch * BYTES_PER_LOOKUP
    achieved = bytes_per_launch / (ms_per_step / 1e3) / 1e9
    traffic, traffic_note = load_traffic(ms_per_step, lookups_per_launch)
    print(f"# {args.steps} frames: {ms:.1f} ms on HIP events, {wall:.2f} s wall; "
          f"{lookups_per_launch:.3e} lookups/frame, {hit_samples / args.steps:.4g} "
          f"volume-hitting samples/frame", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(wl, args.cpu_seconds)
    nrc = nrc_v1 = None
    if not args.no_nrc:
        # radiance cache vs plain on the nanovdb cloud: 1024^3 index space,
        # ~5% active, 8^3 leaves; paths end in the cache at their second
        # volume vertex, 65,536 training samples per frame
        t0 = time.perf_counter()
        wl3 = wls.cloud_nanovdb_pathtrace(n=args.ttr_size, width=args.width, height=args.height,
                                          phase_g=PHASE_G)
        print(f"# generated {wl3.name} in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
              flush=True)
        ref3 = reference_scene(pkg, ctx, wl3)
        nrc = nrc_vs_plain(pkg, ctx, wl3, ref=ref3, oracle_stride=0 if args.no_cpu else 24,
                           query_vertex=args.nrc_query_vertex, train_samples=65536,
                           train_extra=args.nrc_train_extra)
        print(f"# time to relMSE 1e-2: {json.dumps(nrc)}", file=sys.stderr, flush=True)
        # the same with queries at the first volume vertex (32,768 samples)
        nrc_v1 = nrc_vs_plain(pkg, ctx, wl3, ref=ref3, oracle_stride=0, query_vertex=1,
                              train_samples=32768, train_extra=args.nrc_train_extra)
        print(f"# time to relMSE 1e-2, first-vertex queries: {json.dumps(nrc_v1)}",
              file=sys.stderr, flush=True)
        del ref3

    other = {"raycast_blobs": raycast_frame_time(pkg, ctx, cpu=not args.no_cpu)}
    print(f"# raycast: {json.dumps(other['raycast_blobs'])}", file=sys.stderr, flush=True)
    other["raycast_blobs512_1080p"] = raycast_frame_time(
        pkg, ctx, n=512, width=1920, height=1080, frames=8, warmup=2, cpu=False,
        note="512 MiB volume, larger than the 256 MiB MALL; algorithmic bytes count 32 B per "
             "lookup, cache hits included")
    print(f"# raycast 512: {json.dumps(other['raycast_blobs512_1080p'])}", file=sys.stderr,
          flush=True)
    if not args.no_mixed:
        other["mixed_surfaces_volume"] = mixed_frame_time(pkg, ctx, args.width, args.height)
        print(f"# mixed scene: {json.dumps(other['mixed_surfaces_volume'])}", file=sys.stderr,
              flush=True)
        other["mixed_surfaces_volume_lbvh"] = mixed_frame_time(pkg, ctx, args.width,
                                                               args.height, builder="lbvh")
        print(f"# mixed scene, GPU-built BVH: "
              f"{json.dumps(other['mixed_surfaces_volume_lbvh'])}", file=sys.stderr, flush=True)
    if not args.no_nrc and not args.no_c4:
        # unstructured-plus-cache: 2,058,000 tetrahedra, 3840x2160, plain vs
        # the cache in fp16 and bf16, queries at the second volume vertex
        t0 = time.perf_counter()
        wl4 = wls.tets_pathtrace()
        gs4 = wls.GpuScene(ctx, wl4)
        inf = gs4.volume.field.info
        setup = time.perf_counter() - t0
        c4 = nrc_vs_plain(pkg, ctx, wl4, volume=gs4.volume, oracle_stride=0, ref_frames=1024,
                          precisions=("fp16", "bf16"), query_vertex=args.nrc_query_vertex,
                          train_samples=65536, train_extra=args.nrc_train_extra)
        c4["field"] = {"tetrahedra": int(inf.num_cells), "grid_dims": list(inf.grid_dims),
                       "mean_list": round(inf.mean_list, 2),
                       "expected_tests": round(inf.expected_tests, 3),
                       "build_ms": round(inf.build_ms, 1),
                       "device_MB": round(inf.device_bytes / 1e6, 1), "setup_s": round(setup, 1)}
        other["tets_time_to_relmse"] = c4
        print(f"# tets: {json.dumps(c4)}", file=sys.stderr, flush=True)

    mfma = mfma_rates(pkg, ctx)
    print(f"# cache MFMA kernels: {json.dumps(mfma)}", file=sys.stderr, flush=True)

    out = {
        "metric": "Msamples/s at 1080p; time-to-relMSE 1e-2, NRC vs plain pathtrace",
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"pathtrace fBm {args.size}^3 f32, density x50, 32^3 majorants, "
                        f"albedo 0.8, Henyey-Greenstein g={PHASE_G}, directional + ambient, "
                        f"<=8 bounces, {args.width}x{args.height}, 1 spp/frame; camera at "
                        f"(1.5, 1, 2) looking at the volume centre, fovy 40 deg: "
                        f"{hit_samples / samples:.1%} of the samples hit the volume",
            "renderer": "pathtrace",
            "image": [args.width, args.height],
            "volume": [args.size] * 3,
            "spp_per_step": 1,
        },
        "volume_samples_per_s_M": round(hit_samples / (ms / 1e3) / 1e6, 3),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_note": traffic_note,
            "kernel": "k_pathtrace",
            "lookups_per_launch": lookups_per_launch,
            "bytes_per_lookup": BYTES_PER_LOOKUP,
            "lookups_per_volume_sample": round(lookups / max(1, hit_samples), 3),
            "limiter": "the HBM roofline is the one the volume fetch is measured against; the "
                       "kernel runs far below it, bound by dependent fetch latency and VALU "
                       "issue of divergent paths (see traffic: HBM bytes actually moved)",
        },
        "mfma": {"bound": "mfma", "unit": "TFLOP/s", "kernels": "k_nrc_infer, k_nrc_grad",
                 **mfma},
        "cpu_baseline": cpu,
        "time_to_relmse": nrc,
        "time_to_relmse_vertex1": nrc_v1,
        "other_configs": other or None,
    }
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
