"""Benchmark of the erosion hot path on one MI355X.

A step is one map-cycle -- World::erode(N) (world.hpp:36-77) -- of the
1024x1024 workload: periodic two-layer map (fBm + 8 pyramids + 0.05
sediment), 1,000,000 particles per cycle, step-synchronous windows (one
fly() call per particle per window, 514 windows + a closing relax per
cycle), flow tracking on.  Inputs are resident in HBM before the timed
region; the timed region is K whole cycles of the library's default path
(each cycle one HIP-graph launch) bracketed by device synchronisation.

Prints one JSON line with value = particle-steps/s, the roofline of the
tile step kernel (k_step) -- algorithmic bytes per launch over its average
launch duration, measured with HIP events on the library's stream around
every 16th step launch of up to 20 cycles run right after the timed
region -- and the CPU baseline (the reference's sequential algorithm, one
core) timed on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_per_cycle(p, steps, cells):
    """Algorithmic HBM bytes of one map-cycle of the plain method.

    Per particle-step 44 B: the 3x3 float32 snapshot neighbourhood read by
    the normal and the cascade (36 B) and one int64 pickup delta (8 B).  Per
    cell per cycle 36 B: read and zero the int64 delta (16 B), read-modify-
    write the int64 master (16 B), write the float32 snapshot (4 B).  Flow
    tracking adds 32 B per step (four int64 tracks) and 96 B per cell per
    cycle.  The sediment layer's own bytes are not counted.
    """
    per_step = 44 + (32 if p.track else 0)
    per_cell = 36 + (96 if p.track else 0)
    return per_step * steps + per_cell * cells, per_step, per_cell


def measured_traffic(workload, p):
    """HBM bytes per k_step launch from the committed rocprofv3 summary
    (FETCH_SIZE and WRITE_SIZE scaled by a plain copy's reading), when it
    covers this configuration."""
    path = os.path.join(ROOT, "profiles", "k_step_%s_hbm_traffic.txt" % workload)
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for ln in f:
            if ":" in ln and not ln.startswith("#"):
                k, v = ln.split(":", 1)
                vals[k.strip()] = v.strip()
    cfg = "window=%d layers=%d track=%d" % (p.window, p.layers, int(p.track))
    if vals.get("config") != cfg or "traffic_bytes_per_launch" not in vals:
        return None
    return float(vals["traffic_bytes_per_launch"])


class pinned_to_one_core:
    """Pin the calling thread (the one that runs the sequential port) to the
    first core it may use, and restore its affinity afterwards."""

    def __enter__(self):
        self.saved = os.sched_getaffinity(0)
        self.core = min(self.saved)
        os.sched_setaffinity(0, {self.core})
        return self

    def __exit__(self, *exc):
        os.sched_setaffinity(0, self.saved)


def _seq_run(O, p, w, h, s, particles, cycles, seed):
    import numpy as np

    hh, ss = h.ravel().copy(), None if s is None else s.ravel().copy()
    ema = np.zeros(4 * h.size, np.float32)
    t0 = time.perf_counter()
    steps = O.erode_seq(p, w["nx"], w["ny"], hh, ss, ema, particles=particles, n_cycles=cycles, seed=seed)
    return steps, time.perf_counter() - t0


def cpu_baseline(workload, seed, target_s, full=False):
    """The reference's sequential algorithm (oracle port) on one pinned core:
    by default a bounded sample of the workload's first cycle (~target_s
    seconds); with full=True one whole cycle of it, the 256^2 CPU workload in
    full (100 cycles x 2,000 particles), and the batched oracle on every core
    this process may use, for information."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from simplewinderosion_amd import workloads

    w = workloads.WORKLOADS[workload]
    h, s = workloads.make_inputs(workload, seed)
    p = O.params_from(workloads.params(workload, seed))
    cores_all = len(os.sched_getaffinity(0))
    with pinned_to_one_core() as pin:
        if not full:
            # calibrate on a small sample, then run ~target_s seconds of particles
            steps, dt = _seq_run(O, p, w, h, s, 200, 1, seed)
            n = max(200, int(target_s * steps / dt / (p.max_age + 1)))
            steps, dt = _seq_run(O, p, w, h, s, n, 1, seed)
            return {"value": steps / dt, "unit": "particle-steps/s", "cores": 1, "kind": "port",
                    "sample": "%d particles (%d particle-steps) of one %dx%d cycle, sequential glibc-rand order, "
                              "%.1f s on core %d" % (n, steps, w["nx"], w["ny"], dt, pin.core)}
        steps, dt = _seq_run(O, p, w, h, s, w["particles"], 1, seed)
        print("bench: CPU full cycle done, %.1f s" % dt, file=sys.stderr, flush=True)
        out = {"value": steps / dt, "unit": "particle-steps/s", "cores": 1, "kind": "port",
               "sample": "one full %dx%d cycle: %d particles, %d particle-steps, sequential glibc-rand order, "
                         "%.1f s on core %d" % (w["nx"], w["ny"], w["particles"], steps, dt, pin.core)}
        a = workloads.WORKLOADS["cpu256"]
        ha, sa = workloads.make_inputs("cpu256", seed)
        pa = O.params_from(workloads.params("cpu256", seed))
        steps_a, dt_a = _seq_run(O, pa, a, ha, sa, a["particles"], a["cycles"], seed)
        out["cpu256_full"] = {"value": steps_a / dt_a, "seconds": dt_a, "particle_steps": steps_a, "cores": 1,
                              "sample": "256x256, %d cycles x %d particles" % (a["cycles"], a["particles"])}
        print("bench: CPU 256x256 workload done, %.1f s" % dt_a, file=sys.stderr, flush=True)
    # the batched oracle (OpenMP over particles) on all cores, one cycle
    st = O.OracleState(w["nx"], w["ny"], h, sediment=s, params=p)
    t0 = time.perf_counter()
    st.erode(w["particles"], 1, nthreads=cores_all)
    dt_b = time.perf_counter() - t0
    out["batched_oracle"] = {"value": st.steps.value / dt_b,
                             "seconds": dt_b, "cores": cores_all,
                             "sample": "one full %dx%d cycle, batched-synchronous method" % (w["nx"], w["ny"])}
    return out


def calibrate_snap(side):
    """A second kernel of known traffic for the counter scale: the snapshot
    conversion of a side x side map's int64 master (k_snap: 8 bytes read and
    4 written per cell), run three times by uploading the master."""
    import numpy as np

    from simplewinderosion_amd import World

    w = World(side, side, periodic=True)
    m = np.zeros(side * side, np.int64)
    for _ in range(3):
        w.upload_master(m)
    w.synchronize()
    w.close()
    print(json.dumps({"snap_cells": side * side, "snap_bytes_read": 8 * side * side,
                      "snap_bytes_written": 4 * side * side, "launches": 3}), flush=True)


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


def apply_tuning(target, args, particles):
    """Launch knobs from the command line, or the autotuner's choice
    (--autotune); returns the setting in effect."""
    from simplewinderosion_amd import capi

    res = None
    if args.autotune:
        res = target.autotune(particles, trial_cycles=2, cache=None)
    for knob, v in ((capi.SWE_TUNE_SHAPE, args.shape), (capi.SWE_TUNE_TILE_ORDER, args.tile_order),
                    (capi.SWE_TUNE_GRAPH, args.graph), (capi.SWE_TUNE_HEAVY_FIRST, args.heavy_first)):
        if v is not None:
            target.set_tune(knob, v)
    out = {"autotuned": bool(args.autotune),
           "shape": target.get_tune(capi.SWE_TUNE_SHAPE), "tile_order": target.get_tune(capi.SWE_TUNE_TILE_ORDER),
           "graph": target.get_tune(capi.SWE_TUNE_GRAPH), "heavy_first": target.get_tune(capi.SWE_TUNE_HEAVY_FIRST)}
    if res is not None:
        out["candidates"] = res["candidates"]
        out["tuned_ms_per_cycle"] = res["ms_per_cycle"]
    return out


def bench_surrogate(args):
    """The learned surrogate end to end on one GPU: training pairs generated
    by the plain method (64 maps of 256^2 x 400 cycles, random winds, K =
    50), training under the map's symmetries, then a rollout of K = 50
    blocks on the 1024^2 map with its 8
    obstacles, the plain method judging every 10th block (blocks within 1.5x
    the plain method's own run-to-run RMS from the same state accepted; that
    RMS measured at every judged state by a second plain run) and taking
    over after a rejection.  value = map-cycles/s of the whole rollout (surrogate blocks,
    judges and fallback blocks); the plain method's rate beside it."""
    import math

    import torch

    from simplewinderosion_amd import surrogate as S
    from simplewinderosion_amd import workloads

    def log(msg):
        print("bench: " + msg, file=sys.stderr, flush=True)

    K = 50
    states, az, t_gen = S.generate_states(args.sur_terrains, args.sur_cycles, seed=args.seed,
                                          workload=args.sur_train_workload, log=log)
    model = S.Surrogate(seed=args.seed)
    S.fit_scales(model, states, K)
    curve, t_train = S.train(model, states, az, K, args.sur_steps, batch=8, lr=1e-3, log=log, log_every=250,
                             pushforward=args.sur_pushforward, pushforward_blocks=args.sur_pushforward_blocks,
                             graph=args.sur_graph, crop=args.sur_crop, border=args.sur_border)
    side = args.sur_crop if args.sur_crop else states.shape[-1]
    fps = 3 * S.Surrogate.flops_forward(8, side, side)
    del states
    torch.cuda.empty_cache()
    world = workloads.make_world("surr1024", seed=7, deposit_on_death=True)
    plain = workloads.make_world("surr1024", seed=7, deposit_on_death=True)
    ps = workloads.WORKLOADS["surr1024"]["params"]["pspeed"]
    n = workloads.WORKLOADS["surr1024"]["particles"]
    rec = S.rollout(model, world, plain, n, args.sur_blocks, math.atan2(ps[2], ps[0]), judge_every=10,
                    tol_rms=1e-3, tol_noise=1.5, noise_every_judge=not args.sur_first_floor, log=log)
    world.close()
    plain.close()
    rate = args.sur_blocks * K / rec["t_total_s"]
    # the plain method's rate comes from the judged blocks (none in a run shorter than judge_every)
    plain_rate = K / rec["plain_seconds_per_block"] if rec["plain_seconds_per_block"] > 0 else None
    judged = rec["judged"]
    line = {"metric": "map-cycles/s", "value": rate, "unit": "map-cycles/s", "n_gpus": 1,
            "steps": args.sur_blocks, "warmup": 0, "ms_per_step": 1e3 * rec["t_total_s"] / args.sur_blocks,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "bf16 MFMA (fp32 accumulation); plain method f32 with int64 fixed point",
            "data": "synthetic", "config": {"workload": "surrogate rollout 1024x1024 + 8 pyramids", "K": K,
                                            "blocks": args.sur_blocks, "judge_every": 10,
                                            "parallelism": "single GPU"},
            "plain_map_cycles_per_s": plain_rate, "speedup_vs_plain": rate / plain_rate if plain_rate else None,
            "blocks": {"surrogate": rec["surrogate_blocks"], "plain_fallback": rec["plain_blocks"],
                       "judged": rec["judge_blocks"],
                       "accepted": sum(1 for j in judged if j["accepted"])},
            "judged_rms": [round(j["rms_err"], 6) for j in judged],
            "no_change_rms": [round(j["rms_change_plain"], 6) for j in judged],
            "noise_floor_rms": [round(j["noise_floor_rms"], 6) if "noise_floor_rms" in j else None for j in judged],
            "tolerance_rms": [round(j["tol_rms"], 6) for j in judged],
            "noise_floor_at": "each judged state" if not args.sur_first_floor else "the first judged state",
            "mass_drift_fx": rec["mass_drift_fx"], "mass_exact": rec["mass_exact"],
            "training": {"workload": args.sur_train_workload, "crop": args.sur_crop, "border": args.sur_border,
                         "maps": args.sur_terrains, "cycles_per_map": args.sur_cycles, "steps": args.sur_steps,
                         "batch": 8, "graph": args.sur_graph, "pushforward": args.sur_pushforward,
                         "pushforward_blocks": args.sur_pushforward_blocks,
                         "final_loss": curve[-1][1], "tflops": fps * args.sur_steps / t_train / 1e12,
                         "bf16_dense_peak_tflops": 2500.0, "data_erode_seconds": t_gen}}
    print(json.dumps(line), flush=True)


def bench_surrogate_layers(args):
    """The surrogate's convolution kernels layer by layer at the training
    shape (batch 8, 256^2 network): for every layer the forward (bias +
    ReLU), the data gradient (flipped weights, the ReLU mask of the layer
    below) and the weight and bias gradients, each timed over 20 calls with
    events on the stream the kernels run on, from seeded synthetic tensors.
    Prints one JSON line: TFLOP/s of the three passes over all layers
    (3 x sum 2*Cin*Cout*9*pixels) against the BF16 dense MFMA peak."""
    import torch
    from simplewinderosion_amd import nn, surrogate as S

    reps = 20

    def timeit(fn):
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) * 1000.0 / reps

    layers, flops, total = [], 0.0, 0.0
    for name, ci, co, _ in S.LAYERS:
        N, H = 8, 256 >> S.LEVEL[name]
        g = torch.Generator(device="cpu").manual_seed(args.seed)
        x = torch.randn((N, H, H, ci), generator=g).to("cuda").to(torch.bfloat16)
        w = (torch.randn((co, 9, ci), generator=g) * 0.1).to("cuda").to(torch.bfloat16)
        b = torch.randn((co,), generator=g).to("cuda")
        dy = (torch.randn((N, H, H, co), generator=g) * 0.01).to("cuda").to(torch.bfloat16)
        y = torch.empty((N, H, H, co), device="cuda", dtype=torch.bfloat16)
        dx = torch.empty((N, H, H, ci), device="cuda", dtype=torch.bfloat16)
        wt = nn.flip(w)
        nf, nd = nn.conv_work(N, H, H, ci, co), nn.conv_work(N, H, H, co, ci)
        wsf = torch.empty(nf, device="cuda") if nf else None
        wsd = torch.empty(nd, device="cuda") if nd else None
        dw = torch.empty((co, 9, ci), device="cuda")
        db = torch.empty((co,), device="cuda")
        work = nn.wgrad(x, dy, dw, db)
        t_f = timeit(lambda: nn.conv3x3(x, w, b, relu=True, out16=y, work=wsf))
        # the first layer's input gradient is never needed in training
        t_d = timeit(lambda: nn.conv3x3_masked(dy, wt, x, dx, work=wsd)) if name != "enc0a" else 0.0
        t_w = timeit(lambda: nn.wgrad(x, dy, dw, db, work))
        f = 2.0 * ci * co * 9 * N * H * H
        flops += f * (3 if name != "enc0a" else 2)
        total += t_f + t_d + t_w
        layers.append({"layer": name, "cin": ci, "cout": co, "side": H, "forward_us": round(t_f, 1),
                       "data_grad_us": round(t_d, 1), "weight_grad_us": round(t_w, 1)})
    tflops = flops / (total * 1e-6) / 1e12
    print(json.dumps({
        "metric": "TFLOP/s", "value": tflops, "unit": "TFLOP/s", "n_gpus": 1, "steps": reps, "warmup": 1,
        "ms_per_step": total / 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 MFMA (fp32 accumulation)", "data": "synthetic",
        "config": {"workload": "surrogate layers, batch 8 at 256x256", "layers": len(layers),
                   "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": tflops, "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": tflops / 2500.0, "traffic": None},
        "layers": layers}))


def bench_batch(args):
    """The multi-terrain workload: 64 terrains (sides log-uniform in
    128..1024, own seed, wind azimuth uniform at 45 degrees elevation,
    particles per cycle = area) eroded by one launch per window."""
    from simplewinderosion_amd import workloads

    terrains = workloads.batch64(seed=args.seed)
    b = workloads.make_batch(terrains)
    parts = [t["particles"] for t in terrains]
    tuned = apply_tuning(b, args, parts)
    print("bench: batch64, %d terrains, %d cells, warmup %d, steps %d"
          % (len(terrains), sum(parts), args.warmup, args.steps), file=sys.stderr, flush=True)
    b.erode(parts, n_cycles=args.warmup)
    b.synchronize()
    b.reset_stats()
    t0 = time.perf_counter()
    b.erode(parts, n_cycles=args.steps)
    b.synchronize()
    el = time.perf_counter() - t0
    st = b.stats()
    line = {"metric": "particle-steps/s", "value": st["particle_steps"] / el, "unit": "particle-steps/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * el / args.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32 (int64 fixed-point accumulation)", "data": "synthetic",
            "config": {"workload": "batch64", "terrains": len(terrains), "cells": sum(parts),
                       "particles_per_cycle": sum(parts), "window": 1, "layers": 2, "periodic": 1,
                       "parallelism": "single GPU"},
            "tiles": {"max_particles_per_tile": st["max_tile_particles"]}, "tuning": tuned}
    b.close()
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="gpu1024")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=25.0)
    ap.add_argument("--cpu-full", action="store_true",
                    help="CPU baseline on one whole cycle, plus the 256^2 CPU workload in full and the batched "
                         "oracle on all cores (several minutes)")
    ap.add_argument("--window", type=int, default=None)
    ap.add_argument("--cascade-order", type=int, default=None,
                    help="-1 automatic, 0 synchronous, 1 colour-ordered cascade (default: the workload's)")
    ap.add_argument("--track", type=int, default=None)
    ap.add_argument("--tile-order", type=int, default=None,
                    help="SWE_TUNE_TILE_ORDER for the step kernel (default: the library's)")
    ap.add_argument("--graph", type=int, default=None,
                    help="SWE_TUNE_GRAPH: 1 = a cycle is one HIP graph launch (used when timing is off)")
    ap.add_argument("--heavy-first", type=int, default=None,
                    help="SWE_TUNE_HEAVY_FIRST: 1 = tiles crowded in one window run first in the next, "
                         "2 = and light ones last (the library's default)")
    ap.add_argument("--shape", type=int, default=None,
                    help="SWE_TUNE_SHAPE: the step kernel's compiled tile shape (0 = 16x32 cells, the default)")
    ap.add_argument("--autotune", action="store_true",
                    help="run the library's autotuner (shapes, tile orders, graphs, heavy-first) before timing")
    ap.add_argument("--sur-terrains", type=int, default=64, help="surrogate: training maps")
    ap.add_argument("--sur-cycles", type=int, default=400, help="surrogate: plain cycles per training map")
    ap.add_argument("--sur-steps", type=int, default=12000, help="surrogate: Adam steps")
    ap.add_argument("--sur-train-workload", default="surr256",
                    help="surrogate: workload of the training maps (surr256, or surr1024 with --sur-crop)")
    ap.add_argument("--sur-crop", type=int, default=0,
                    help="surrogate: train on crops of this side from larger maps (0: whole maps)")
    ap.add_argument("--sur-border", type=int, default=0,
                    help="surrogate: band along a crop's edges left out of the loss gradient")
    ap.add_argument("--sur-blocks", type=int, default=40, help="surrogate: rollout blocks of K = 50 cycles")
    ap.add_argument("--sur-graph", action="store_true",
                    help="surrogate: replay one captured HIP graph per training step instead of launching its "
                         "kernels one by one (the same bits; measured 4%% slower)")
    ap.add_argument("--sur-first-floor", action="store_true",
                    help="surrogate: measure the plain method's run-to-run RMS at the first judged state only "
                         "(the earlier protocol) instead of at every judged state")
    ap.add_argument("--sur-pushforward", type=float, default=0.0,
                    help="surrogate: share of the later training batches started from the network's own rollout")
    ap.add_argument("--sur-pushforward-blocks", type=int, default=1,
                    help="surrogate: longest own rollout (in blocks of K cycles) a pushforward batch starts from")
    ap.add_argument("--timing-stride", type=int, default=16,
                    help="HIP events around every n-th step launch of the sampling cycles run after the timed "
                         "region, for the roofline (0: no sampling)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (counter calibration)")
    ap.add_argument("--calibrate-snap", type=int, default=0, metavar="SIDE",
                    help="only run the snapshot conversion of a SIDE x SIDE map (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.calibrate_snap:
        calibrate_snap(args.calibrate_snap)
        return

    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (a batch would shard by "
              "terrain; one map by bands of tiles that exchange snapshot halos and crossing particles "
              "with their two neighbours after each window)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch64":
        bench_batch(args)
        return
    if args.workload == "surrogate":
        bench_surrogate(args)
        return
    if args.workload == "surrogate-layers":
        bench_surrogate_layers(args)
        return

    from simplewinderosion_amd import workloads

    wl = workloads.WORKLOADS[args.workload]
    over = {}
    if args.window is not None:
        over["window"] = args.window
    if args.track is not None:
        over["track"] = bool(args.track)
    if args.cascade_order is not None:
        over["cascade_order"] = args.cascade_order
    world = workloads.make_world(args.workload, seed=args.seed, **over)
    p = world.params()
    tuned = apply_tuning(world, args, [wl["particles"]])
    N = wl["particles"]
    R = p.window if p.window > 0 else p.max_age + 2
    steps_per_cycle = N * (p.max_age + 1)

    mass0 = world.audit()["height_sum"]
    print("bench: %s %dx%d, %d particles/cycle, window %d, layers %d, track %d; warmup %d, steps %d"
          % (args.workload, world.nx, world.ny, N, R, p.layers, p.track, args.warmup, args.steps),
          file=sys.stderr, flush=True)
    world.erode(N, n_cycles=args.warmup)
    world.synchronize()
    # the timed region runs the library's default path (a map-cycle replayed
    # as one HIP graph; no events inside)
    world.set_timing(0)
    world.reset_stats()
    world.synchronize()
    t0 = time.perf_counter()
    world.erode(N, n_cycles=args.steps)
    world.synchronize()
    t1 = time.perf_counter()
    st = world.stats()
    elapsed = t1 - t0
    total_steps = st["particle_steps"]  # counted on the device
    # on a periodic map every particle makes max_age + 1 steps (a bounded one
    # loses those that leave it)
    if p.periodic and total_steps != steps_per_cycle * args.steps:
        raise SystemExit("device step count %d != %d particles x %d steps x %d cycles"
                         % (total_steps, N, p.max_age + 1, args.steps))
    steps_per_cycle = total_steps / args.steps
    value = total_steps / elapsed

    # step kernel launch durations: HIP events on the library's stream around
    # every timing_stride-th step launch, over cycles run right after the
    # timed region (events need individual launches, so these cycles run
    # outside the graph); their own rate is reported beside the timed one
    sample_cycles = max(2, min(args.steps, 20))
    sampled_ms = None
    avg_launch_ms = None
    if args.timing_stride > 0:
        world.set_timing(args.timing_stride)
        world.reset_stats()
        world.synchronize()
        ts0 = time.perf_counter()
        world.erode(N, n_cycles=sample_cycles)
        world.synchronize()
        sampled_ms = 1e3 * (time.perf_counter() - ts0) / sample_cycles
        st2 = world.stats()
        avg_launch_ms = st2["march_ms"] / st2["march_launches"] if st2["march_launches"] else None
        world.set_timing(0)
    else:
        st2 = {"march_launches": 0}
    audit = world.audit()

    # step kernel: algorithmic bytes per launch / HIP-event launch time
    launches_per_cycle = (p.max_age + 2) + 1 if R == 1 else None
    cyc_bytes, per_step, per_cell = algorithmic_bytes_per_cycle(p, steps_per_cycle, world.nx * world.ny)
    achieved = None
    bytes_per_launch = None
    if avg_launch_ms and launches_per_cycle:
        bytes_per_launch = cyc_bytes / launches_per_cycle
        achieved = bytes_per_launch / (avg_launch_ms / 1e3) / 1e9
    traffic = measured_traffic(args.workload, p)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": traffic,
                "kernel": "k_step", "algorithmic_bytes_per_launch": bytes_per_launch,
                "bytes_per_step": per_step, "bytes_per_cell_cycle": per_cell,
                "launches_per_cycle": launches_per_cycle, "avg_launch_ms": avg_launch_ms,
                "timed_launches": st2["march_launches"], "timing_stride": args.timing_stride,
                "sampled_cycles": sample_cycles if args.timing_stride > 0 else 0,
                "sampled_ms_per_cycle": sampled_ms}
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
                   "cascade": "colour-ordered" if (p.cascade_order == 1 or (p.cascade_order < 0 and p.settling > 0.25))
                   else "synchronous",
                   "periodic": int(p.periodic), "parallelism": "single GPU"},
        "map_cycles_per_s": args.steps / elapsed,
        "mass_exact": None,
        "tiles": {"max_particles_per_tile": st["max_tile_particles"],
                  "crowded_tile_windows": st["crowded_tile_windows"], "far_moves": st["far_particles"],
                  "heavy_first_tile_windows": st["heavy_tile_windows"]},
        "roofline": roofline,
        "tuning": tuned,
    }
    line["mass_exact"] = bool(audit["height_sum"] + audit["lost"] == mass0)
    world.close()
    if not args.no_cpu:
        print("bench: CPU baseline (sequential port, 1 core, ~%.0f s)" % args.cpu_seconds, file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(args.workload, args.seed, args.cpu_seconds, full=args.cpu_full)
        line["gpu_over_cpu"] = value / line["cpu_baseline"]["value"]
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
