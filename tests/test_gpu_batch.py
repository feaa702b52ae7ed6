"""Batches: every terrain of a multi-terrain context equals the same terrain
eroded alone (and the oracle), bit for bit."""
import numpy as np
import pytest

import oracle as O
from simplewinderosion_amd import capi
from simplewinderosion_amd import Batch, World, Wind, glm_normalize, workloads

pytestmark = pytest.mark.gpu


def test_batch_terrains_equal_single_runs(statics):
    sizes = [(96, 128), (40, 300), (130, 70), (64, 64), (17, 33)]
    print("case batch", sizes, flush=True)
    Wind.suspension = 1e-3
    World.settling = 0.05
    rng = np.random.default_rng(0)
    winds = [glm_normalize((np.cos(a), 0.3, np.sin(a))) for a in rng.uniform(0, 6.28, len(sizes))]
    seeds = [11, 22, 33, 44, 55]
    heights = [(rng.random(s) * 0.4).astype(np.float32) for s in sizes]
    seds = [np.full(s, 0.004, np.float32) for s in sizes]
    parts = [int(0.8 * a * b) for a, b in sizes]
    b = Batch(sizes, periodic=True, layers=2, abrasion=1e-3)
    for k in range(len(sizes)):
        b.set_terrain(k, seeds[k], winds[k])
        b.upload(k, heights[k], seds[k])
    b.erode(parts, n_cycles=2)
    for k, (nx, ny) in enumerate(sizes):
        World.SEED = seeds[k]
        Wind.pspeed = winds[k]
        w = World(nx, ny, periodic=True, layers=2, abrasion=1e-3)
        w.upload(heights[k], seds[k])
        w.erode(parts[k], n_cycles=2)
        zb, sb = b.master(k)
        zw, sw = w.master()
        assert np.array_equal(zb, zw) and np.array_equal(sb, sw), k
        db, dw = b.download(k), w.download()
        for f in ("massflow", "momentum"):
            assert np.array_equal(db[f], dw[f]), (k, f)
        assert b.audit(k)["lost"] == w.audit()["lost"], k
        if k == 1:
            st = O.OracleState(nx, ny, heights[k], sediment=seds[k], params=O.params_from(w.params()))
            st.erode(parts[k], 2)
            assert np.array_equal(zw.ravel(), st.zm)
        w.close()
    b.close()


def test_batch64_workload_builds_and_runs_one_cycle(statics):
    """The 64-terrain workload (sides log-uniform in 128..1024): one cycle,
    exact mass ledger on every terrain."""
    terr = workloads.batch64(seed=3)
    print("case batch64 cells", sum(t["nx"] * t["ny"] for t in terr), flush=True)
    b = workloads.make_batch(terr)
    assert b.get_tune(capi.SWE_TUNE_HEAVY_FIRST) == 1  # the default above 4096 tiles
    m0 = [b.audit(k)["height_sum"] for k in range(len(terr))]
    b.erode([t["particles"] for t in terr], n_cycles=1)
    b.synchronize()
    for k in range(len(terr)):
        a = b.audit(k)
        assert a["height_sum"] + a["lost"] == m0[k], k
    b.close()
