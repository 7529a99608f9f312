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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run the plain copy kernel over MIB MiB (counter calibration)")
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
    t0 = time.perf_counter()
    wl = wls.fbm_pathtrace(n=args.size, width=args.width, height=args.height,
                           frames=args.steps)
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
    lookups = int(gs.frame.counter_values()[0])

    samples = wl.width * wl.height * args.steps
    value = samples / (ms / 1e3) / 1e6
    ms_per_step = ms / args.steps
    bytes_per_launch = lookups / args.steps * BYTES_PER_LOOKUP
    achieved = bytes_per_launch / (ms_per_step / 1e3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    print(f"# {args.steps} frames: {ms:.1f} ms on HIP events, {wall:.2f} s wall; "
          f"{lookups / args.steps:.3e} lookups/frame", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(wl, args.cpu_seconds)

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
                        f"directional + ambient, <=8 bounces, {args.width}x{args.height}, "
                        f"1 spp/frame",
            "renderer": "pathtrace",
            "image": [args.width, args.height],
            "volume": [args.size] * 3,
            "spp_per_step": 1,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": "k_pathtrace",
            "lookups_per_launch": lookups / args.steps,
            "bytes_per_lookup": BYTES_PER_LOOKUP,
        },
        "cpu_baseline": cpu,
    }
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
