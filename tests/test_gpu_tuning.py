"""The autotuner (swe_autotune through the C-ABI, tuning.py cache): every
candidate is timed from the same saved state and the state is restored, so
tuning changes nothing but speed."""
import numpy as np
import pytest

from simplewinderosion_amd import World, Wind, capi

pytestmark = pytest.mark.gpu


def test_autotune_keeps_state_and_results(statics, tmp_path):
    print("case autotune", flush=True)
    Wind.suspension = 1e-3
    World.settling = 0.02
    World.SEED = 5
    rng = np.random.default_rng(8)
    h = (rng.random((128, 96)) * 0.5).astype(np.float32)
    s = np.full((128, 96), 0.01, np.float32)
    a = World(128, 96, periodic=True, layers=2, window=1, abrasion=1e-3)
    b = World(128, 96, periodic=True, layers=2, window=1, abrasion=1e-3)
    for w in (a, b):
        w.upload(h, s)
        w.erode(5000)
    before = (a.master(), a.download(), a.audit(), a.stats()["particle_steps"], a.cycle)
    cache = str(tmp_path / "tune.json")
    r = a.autotune(5000, trial_cycles=1, cache=cache)
    assert r["candidates"] == 8 and not r["cached"] and r["ms_per_cycle"] > 0
    assert a.get_tune(capi.SWE_TUNE_TILE_ORDER) == r["tile_order"]
    assert a.get_tune(capi.SWE_TUNE_GRAPH) == r["graph"]
    after = (a.master(), a.download(), a.audit(), a.stats()["particle_steps"], a.cycle)
    for x, y in zip(before[0], after[0]):
        assert np.array_equal(x, y)
    for k in before[1]:
        assert np.array_equal(before[1][k], after[1][k]), k
    assert before[2:] == after[2:]
    # the tuned context continues exactly as an untuned one
    a.erode(5000, n_cycles=2)
    b.erode(5000, n_cycles=2)
    for x, y in zip(a.master(), b.master()):
        assert np.array_equal(x, y)
    assert a.audit() == b.audit()
    # a second call with the same workload is served from the cache
    r2 = a.autotune(5000, cache=cache)
    assert r2["cached"] and r2["tile_order"] == r["tile_order"] and r2["graph"] == r["graph"]
    a.close()
    b.close()
