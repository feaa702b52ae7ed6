"""The batched method on the GPU -- synchronous and colour-ordered cascade --
against the reference's sequential algorithm (oracle port, one core per
run), on the 1024x1024 workload's recipe (fBm + 8 pyramids, two layers,
wind normalize(1,0,1), suspension 1e-4, settling 0.01) at one particle per
cell, compared by height distribution, height change, dune wavelength and
mass.  Each method runs from the same map with two particle seeds, so the
seed-to-seed spread of each is the noise floor.  Writes one JSON object.

  python tests/statistical/compare_density.py [--n 1024] [--cycles 2] [--out file]
"""
import argparse
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

import numpy as np  # noqa: E402


def sequential(args):
    """One sequential run (its own process: one core)."""
    n, particles, cycles, seed = args
    import oracle as O
    from simplewinderosion_amd import workloads

    h, s = workloads.make_inputs("gpu1024", 1, nx=n, ny=n)
    p = O.params_from(workloads.params("gpu1024", seed))
    hh, ss = h.ravel().copy(), s.ravel().copy()
    ema = np.zeros(4 * h.size, np.float32)
    t0 = time.perf_counter()
    for c in range(cycles):  # one cycle per call: the glibc stream continues (srand on the first only)
        O.erode_seq(p, n, n, hh, ss, ema, particles=particles, n_cycles=1, seed=seed if c == 0 else None)
        print("sequential seed %d: cycle %d done, %.1f s" % (seed, c + 1, time.perf_counter() - t0), flush=True)
    return hh.reshape(n, n), time.perf_counter() - t0


def batched(n, particles, cycles, seed, cascade_order):
    from simplewinderosion_amd import World, workloads

    w = workloads.make_world("gpu1024", seed=1, nx=n, ny=n, cascade_order=cascade_order)
    World.SEED = seed
    t0 = time.perf_counter()
    w.erode(particles, n_cycles=cycles)
    w.synchronize()
    dt = time.perf_counter() - t0
    h = w.download()["height"].astype(np.float64)
    w.close()
    return h, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--cycles", type=int, default=2)
    ap.add_argument("--particles", type=int, default=None, help="per cycle (default n*n)")
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    from simplewinderosion_amd import stats, workloads

    n, cycles = args.n, args.cycles
    particles = args.particles or n * n
    h, _ = workloads.make_inputs("gpu1024", 1, nx=n, ny=n)
    seeds = (1, 2)
    print("sequential runs: %dx%d, %d particles x %d cycles, seeds %s (one process each)"
          % (n, n, particles, cycles, seeds), flush=True)
    with ProcessPoolExecutor(max_workers=len(seeds)) as ex:
        fut = [ex.submit(sequential, (n, particles, cycles, s)) for s in seeds]
        runs = {"seq": []}
        times = {"seq": []}
        # the GPU runs meanwhile
        for name, order in (("sync", 0), ("colour", 1)):
            runs[name], times[name] = [], []
            for s in seeds:
                hb, dt = batched(n, particles, cycles, s, order)
                runs[name].append(hb)
                times[name].append(dt)
                print("gpu %s seed %d: %.3f s" % (name, s, dt), flush=True)
        for f in fut:
            hs, dt = f.result()
            runs["seq"].append(hs.astype(np.float64))
            times["seq"].append(dt)
            print("sequential: %.1f s" % dt, flush=True)
    res = {"map": "%dx%d" % (n, n), "recipe": "gpu1024", "particles_per_cycle": particles, "cycles": cycles,
           "seconds": times,
           "seq_vs_seq": stats.compare(h, runs["seq"][0], runs["seq"][1])}
    for name in ("sync", "colour"):
        res[name + "_vs_seq"] = [stats.compare(h, runs["seq"][i], runs[name][i]) for i in range(len(seeds))]
        res[name + "_vs_" + name] = stats.compare(h, runs[name][0], runs[name][1])
    line = json.dumps(res)
    print(line, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(line + "\n")


if __name__ == "__main__":
    main()
