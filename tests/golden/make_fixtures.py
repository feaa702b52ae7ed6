"""Writes the golden fixtures of the batched method with the oracle
(oracle/swe_oracle.c): small seeded inputs and the exact int64 masters, EMA
fields, lost mass and step counts after a few cycles.

  python tests/golden/make_fixtures.py

The fixtures pin the oracle against later changes (tests/test_golden.py
re-runs it) and the GPU path against the committed bits
(tests/test_gpu_golden.py).  The reference ships no data of its own for this
path, so these are the project's vectors.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

import oracle as O  # noqa: E402

# name: nx, ny, particles, cycles, params
CASES = {
    "periodic_1layer": (32, 32, 300, 2, dict(periodic=1, window=1, settling=0.05, suspension=0.01, seed=11)),
    "bounded_1layer": (24, 40, 300, 2, dict(periodic=0, window=1, settling=0.05, suspension=0.01, seed=12)),
    "periodic_2layer": (40, 24, 400, 2, dict(periodic=1, window=1, settling=0.02, suspension=0.001, layers=2,
                                              abrasion=0.001, seed=13)),
    "deposit_w4": (32, 32, 300, 2, dict(periodic=1, window=4, settling=0.02, suspension=0.01,
                                        deposit_on_death=1, seed=14)),
    "edge_spawn": (32, 48, 300, 2, dict(periodic=0, window=1, settling=0.05, suspension=0.01, spawn=1, seed=15)),
}


def inputs(name):
    nx, ny, _, _, kw = CASES[name]
    rng = np.random.default_rng(sum(map(ord, name)))
    h = (rng.random((nx, ny)) * 0.5).astype(np.float32)
    s = np.full((nx, ny), 0.01, np.float32) if kw.get("layers", 1) == 2 else None
    return h, s


def run(name):
    nx, ny, particles, cycles, kw = CASES[name]
    h, s = inputs(name)
    st = O.OracleState(nx, ny, h, sediment=s, params=O.reference_params(**kw))
    st.erode(particles, cycles)
    out = {"height_in": h, "zm": st.zm.copy(), "ema": st.ema.copy(), "lost": np.int64(st.lost.value),
           "steps": np.int64(st.steps.value), "cycle": np.uint64(st.cycle.value)}
    if s is not None:
        out["sediment_in"] = s
        out["sm"] = st.sm.copy()
    return out


def main():
    for name in CASES:
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **run(name))
        print("wrote", name)


if __name__ == "__main__":
    main()
