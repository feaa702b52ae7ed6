"""Benchmark of the erosion hot path on one MI355X.

A step is one map-cycle -- World::erode(N) (world.hpp:36-77) -- of the
1024x1024 workload: periodic two-layer map (fBm + 8 pyramids + 0.05
sediment), 1,000,000 particles per cycle, 16-call snapshot windows.  Inputs
are resident in HBM before the timed region; the timed region is K whole
cycles bracketed by device synchronisation.

Prints one JSON line with value = particle-steps/s, the roofline of the
particle-march kernel measured with HIP events on the library's stream, and
the CPU baseline (the reference's sequential algorithm, one core) timed on
a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def march_bytes_per_step(p, R):
    """Algorithmic bytes the march kernel moves per particle-step.

    5-point float32 stencil read (20 B) + int64 pickup delta (8 B) + uint32
    cascade-event count (4 B); sediment layer: + float32 sediment read (4 B)
    + int64 sediment delta (8 B); tracking: + four int64 tracks (32 B);
    particle state load + store once per window: 2 x 44 B / R.
    """
    b = 20 + 8 + 4
    if p.layers == 2:
        b += 4 + 8
    if p.track:
        b += 32
    b += 2 * 44.0 / R
    return b


def cpu_baseline(workload, seed, target_s):
    """The reference's sequential algorithm (oracle port), one core, on a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    from simplewinderosion_amd import workloads

    w = workloads.WORKLOADS[workload]
    h, s = workloads.make_inputs(workload, seed)
    p = O.params_from(workloads.params(workload, seed))
    ema = np.zeros(4 * h.size, np.float32)
    # calibrate on a small sample, then run ~target_s seconds of particles
    hh, ss = h.ravel().copy(), None if s is None else s.ravel().copy()
    t0 = time.perf_counter()
    steps = O.erode_seq(p, w["nx"], w["ny"], hh, ss, ema, particles=200, n_cycles=1, seed=seed)
    dt = time.perf_counter() - t0
    rate = steps / dt
    n = max(200, int(target_s * rate / (p.max_age + 1)))
    hh, ss = h.ravel().copy(), None if s is None else s.ravel().copy()
    t0 = time.perf_counter()
    steps = O.erode_seq(p, w["nx"], w["ny"], hh, ss, ema, particles=n, n_cycles=1, seed=seed)
    dt = time.perf_counter() - t0
    return {"value": steps / dt, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": "%d particles (%d particle-steps) of one %dx%d cycle, sequential glibc-rand order, %.1f s"
                      % (n, steps, w["nx"], w["ny"], dt)}


def calibrate_copy(mib):
    """A plain device-to-device copy of known size (hipMemcpy, one copy kernel
    per call): the reference point that scales rocprof's FETCH_SIZE /
    WRITE_SIZE counters to bytes."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    n = int(mib) << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)):
        raise SystemExit("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(n))
    for _ in range(3):
        if hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3):  # hipMemcpyDeviceToDevice
            raise SystemExit("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "copies": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="gpu1024")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--window", type=int, default=None)
    ap.add_argument("--track", type=int, default=None)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return

    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the map would shard by "
              "terrain or by particles with an int64 all-reduce of the delta map; see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    from simplewinderosion_amd import workloads

    wl = workloads.WORKLOADS[args.workload]
    over = {}
    if args.window is not None:
        over["window"] = args.window
    if args.track is not None:
        over["track"] = bool(args.track)
    world = workloads.make_world(args.workload, seed=args.seed, **over)
    p = world.params()
    if not p.periodic:
        raise SystemExit("bench counts particle-steps analytically: periodic workloads only")
    N = wl["particles"]
    R = p.window if p.window > 0 else p.max_age + 2
    steps_per_cycle = N * (p.max_age + 1)

    mass0 = world.audit()["height_sum"]
    print("bench: %s %dx%d, %d particles/cycle, window %d, layers %d, track %d; warmup %d, steps %d"
          % (args.workload, world.nx, world.ny, N, R, p.layers, p.track, args.warmup, args.steps),
          file=sys.stderr, flush=True)
    world.erode(N, n_cycles=args.warmup)
    world.synchronize()
    world.set_timing(True)
    world.reset_stats()
    world.synchronize()
    t0 = time.perf_counter()
    world.erode(N, n_cycles=args.steps)
    world.synchronize()
    t1 = time.perf_counter()
    st = world.stats()
    audit = world.audit()
    elapsed = t1 - t0
    total_steps = steps_per_cycle * args.steps
    value = total_steps / elapsed

    # march kernel: algorithmic bytes / HIP-event kernel time
    bps = march_bytes_per_step(p, R)
    march_s = st["march_ms"] / 1e3
    achieved = bps * total_steps / march_s / 1e9 if march_s > 0 else None
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": None,
                "kernel": "k_march", "bytes_per_step": bps,
                "avg_launch_ms": st["march_ms"] / max(1, st["march_launches"]),
                "relax_avg_launch_ms": st["relax_ms"] / max(1, st["relax_launches"])}
    line = {
        "metric": "particle-steps/s",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (int64 fixed-point accumulation)",
        "data": "synthetic",
        "config": {"workload": args.workload, "map": "%dx%d" % (world.nx, world.ny),
                   "particles_per_cycle": N, "window": R, "layers": int(p.layers), "track": int(p.track),
                   "periodic": int(p.periodic), "parallelism": "single GPU"},
        "map_cycles_per_s": args.steps / elapsed,
        "march_ms_total": st["march_ms"],
        "relax_ms_total": st["relax_ms"],
        "mass_exact": None,
        "roofline": roofline,
    }
    line["mass_exact"] = bool(audit["height_sum"] + audit["lost"] == mass0)
    world.close()
    if not args.no_cpu:
        print("bench: CPU baseline (sequential port, 1 core, ~%.0f s)" % args.cpu_seconds, file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(args.workload, args.seed, args.cpu_seconds)
        line["gpu_over_cpu"] = value / line["cpu_baseline"]["value"]
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
