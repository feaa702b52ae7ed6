"""The batched method on the GPU against the reference's sequential
algorithm (oracle port on the host), compared the way two stochastic runs
can be: the distribution of the height change, its RMS, the dune
wavelength and the mass.  The two-layer 256x256 workload's recipe and
particle density (2,000 per 256^2 per cycle) on a 128x128 map, 40 cycles,
two particle seeds per method: the seed-to-seed spread is the noise floor.

Measured (oracle, same sizes): the colour-ordered cascade sits within the
sequential run's own noise (Wasserstein-1 of the change 1.0e-4 against
1.0e-4; RMS change equal to 1e-4 relative); the synchronous cascade, the
benchmark's, is ~3.3x the noise with an RMS change 0.6% higher."""
import numpy as np
import pytest

import oracle as O
from simplewinderosion_amd import World, stats, workloads

pytestmark = pytest.mark.gpu

N, PARTICLES, CYCLES = 128, 500, 40  # 500 / 128^2 = 2000 / 256^2
SEEDS = (1, 2)


@pytest.fixture(scope="module")
def runs():
    h, s = workloads.make_inputs("cpu256", 1, nx=N, ny=N)
    out = {"h": h, "seq": [], "sync": [], "colour": []}
    for seed in SEEDS:
        print("case statistics seed %d" % seed, flush=True)
        p = O.params_from(workloads.params("cpu256", seed))
        hh, ss = h.ravel().copy(), s.ravel().copy()
        ema = np.zeros(4 * h.size, np.float32)
        O.erode_seq(p, N, N, hh, ss, ema, particles=PARTICLES, n_cycles=CYCLES, seed=seed)
        out["seq"].append(hh.reshape(N, N).astype(np.float64))
        for name, order in (("sync", 0), ("colour", 1)):
            w = workloads.make_world("cpu256", seed=1, nx=N, ny=N, cascade_order=order)
            World.SEED = seed
            w.erode(PARTICLES, n_cycles=CYCLES)
            out[name].append(w.download()["height"].astype(np.float64))
            w.close()
    return out


def _w1_change(h, a, b):
    return stats.wasserstein1(a - h, b - h)


def test_colour_cascade_change_within_sequential_noise(statics, runs):
    h = runs["h"]
    noise = max(_w1_change(h, *runs["seq"]), _w1_change(h, *runs["colour"]))
    for i in range(len(SEEDS)):
        d = _w1_change(h, runs["seq"][i], runs["colour"][i])
        assert d <= 1.6 * noise, (i, d, noise)
        rs = float(np.sqrt(np.mean((runs["seq"][i] - h) ** 2)))
        rc = float(np.sqrt(np.mean((runs["colour"][i] - h) ** 2)))
        assert abs(rc - rs) <= 2e-3 * rs, (rs, rc)


def test_synchronous_cascade_gap_is_bounded(statics, runs):
    """The benchmark's synchronous cascade: a measured, bounded departure."""
    h = runs["h"]
    noise = _w1_change(h, *runs["seq"])
    for i in range(len(SEEDS)):
        d = _w1_change(h, runs["seq"][i], runs["sync"][i])
        assert d <= 6 * noise, (i, d, noise)
        rs = float(np.sqrt(np.mean((runs["seq"][i] - h) ** 2)))
        rb = float(np.sqrt(np.mean((runs["sync"][i] - h) ** 2)))
        assert abs(rb - rs) <= 0.02 * rs, (rs, rb)


@pytest.mark.parametrize("method", ["sync", "colour"])
def test_wavelength_and_mass_agree(statics, runs, method):
    h = runs["h"]
    for i in range(len(SEEDS)):
        ws = stats.dune_wavelength(runs["seq"][i] - h)[0]
        wb = stats.dune_wavelength(runs[method][i] - h)[0]
        assert abs(ws - wb) <= 0.25 * ws, (ws, wb)
        noise = stats.relative_mass_difference(*runs["seq"])
        assert stats.relative_mass_difference(runs["seq"][i], runs[method][i]) <= 5 * noise + 1e-7
