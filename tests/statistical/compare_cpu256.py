"""The 256x256 workload (100 cycles x 2,000 particles, two layers, fBm +
0.05 sediment, wind normalize(1,0,1)) run by the batched method on the GPU and
by the reference's sequential algorithm (oracle port) on the host, compared by
height distribution, dune wavelength and mass.  Writes one JSON object.

  python tests/statistical/compare_cpu256.py [--cycles 100] [--out file]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402
from simplewinderosion_amd import World, stats, workloads  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--cycles", type=int, default=100)
    ap.add_argument("--particles", type=int, default=2000)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    nx = workloads.WORKLOADS["cpu256"]["nx"]
    h, s = workloads.make_inputs("cpu256", 1)
    res = {"workload": "cpu256", "cycles": args.cycles, "particles": args.particles}
    for deposit in (0, 1):
        runs = {"seq": [], "gpu": []}
        for seed in (1, 2):
            p = O.params_from(workloads.params("cpu256", seed))
            p.deposit_on_death = deposit
            hh, ss = h.ravel().copy(), s.ravel().copy()
            ema = np.zeros(4 * h.size, np.float32)
            t0 = time.perf_counter()
            steps = O.erode_seq(p, nx, nx, hh, ss, ema, particles=args.particles, n_cycles=args.cycles, seed=seed)
            ts = time.perf_counter() - t0
            runs["seq"].append(hh.reshape(nx, nx))
            # the same initial map (terrain seed 1) for every run; the seed
            # only changes the particles' random streams
            w = workloads.make_world("cpu256", seed=1, deposit_on_death=bool(deposit))
            World.SEED = seed
            t0 = time.perf_counter()
            w.erode(args.particles, n_cycles=args.cycles)
            w.synchronize()
            tg = time.perf_counter() - t0
            runs["gpu"].append(w.download()["height"].astype(np.float64))
            w.close()
            print("deposit %d seed %d: sequential %.1f s (%d steps), gpu %.3f s" % (deposit, seed, ts, steps, tg),
                  flush=True)
        key = "deposit_on_death" if deposit else "reference_loss"
        res[key] = {
            "seq_vs_gpu": stats.compare(h, runs["seq"][0], runs["gpu"][0]),
            "seq_vs_seq": stats.compare(h, runs["seq"][0], runs["seq"][1]),
            "gpu_vs_gpu": stats.compare(h, runs["gpu"][0], runs["gpu"][1]),
        }
    line = json.dumps(res)
    print(line, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(line + "\n")


if __name__ == "__main__":
    main()
