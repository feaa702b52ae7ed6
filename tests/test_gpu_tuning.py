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
    # five shapes, then 4 tile orders x 2 graph modes x 3 heavy-first modes on
    # the fastest (maps of 64 tiles or more) or x 1 below
    assert r["candidates"] in (5 + 8, 5 + 24) and not r["cached"] and r["ms_per_cycle"] > 0
    assert a.get_tune(capi.SWE_TUNE_SHAPE) == r["shape"]
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
    assert r2["shape"] == r["shape"]
    a.close()
    b.close()


@pytest.mark.parametrize("cascade_order", [0, 1])
@pytest.mark.parametrize("periodic", [True, False])
def test_every_tile_shape_gives_the_same_bits(statics, periodic, cascade_order):
    """The step kernel's compiled tile shapes (SWE_TUNE_SHAPE 0..4) on one
    workload with crowded and light tiles, a bounded or periodic map whose
    sides are multiples of no shape, with the synchronous cascade (the step
    kernel's own relax) and the colour-ordered one (its step-only mode):
    identical masters, fields and counters, and the oracle's."""
    import oracle as O

    print("case shapes periodic=%s cascade_order=%d" % (periodic, cascade_order), flush=True)
    Wind.suspension = 1e-3
    World.settling = 0.02
    World.SEED = 9
    nx, ny = 150, 133
    rng = np.random.default_rng(3)
    h = (rng.random((nx, ny)) * 0.5).astype(np.float32)
    s = np.full((nx, ny), 0.01, np.float32)
    results = []
    for shape in range(5):
        w = World(nx, ny, periodic=periodic, layers=2, window=1, abrasion=1e-3, cascade_order=cascade_order)
        w.set_tune(capi.SWE_TUNE_SHAPE, shape)
        assert w.get_tune(capi.SWE_TUNE_SHAPE) == shape
        w.upload(h, s)
        w.erode(30000, n_cycles=2)
        results.append((w.master(), w.download(), w.audit(), w.stats()["particle_steps"]))
        w.close()
    for shape in range(1, 5):
        for x, y in zip(results[0][0], results[shape][0]):
            assert np.array_equal(x, y), shape
        for k in ("massflow", "momentum"):
            assert np.array_equal(results[0][1][k], results[shape][1][k]), (shape, k)
        assert results[0][2:] == results[shape][2:], shape
    from simplewinderosion_amd.world import build_params

    st = O.OracleState(nx, ny, h, sediment=s, params=O.params_from(
        build_params(periodic=periodic, layers=2, window=1, abrasion=1e-3, cascade_order=cascade_order)))
    assert O.load().or_colour_mode(st.p) == cascade_order
    st.erode(30000, 2)
    assert np.array_equal(results[0][0][0].ravel(), st.zm)
    assert np.array_equal(results[0][0][1].ravel(), st.sm)


def test_shape_change_between_cycles_continues_exactly(statics):
    """Switching the shape between cycles of one context changes nothing."""
    print("case shape switch", flush=True)
    Wind.suspension = 1e-3
    World.settling = 0.02
    World.SEED = 4
    rng = np.random.default_rng(5)
    h = (rng.random((96, 160)) * 0.5).astype(np.float32)
    a = World(96, 160, periodic=True, window=1)
    b = World(96, 160, periodic=True, window=1)
    for w in (a, b):
        w.upload(h)
    for shape in (4, 2, 1, 3, 0):
        a.set_tune(capi.SWE_TUNE_SHAPE, shape)
        a.erode(9000)
    b.erode(9000, n_cycles=5)
    for x, y in zip(a.master(), b.master()):
        assert np.array_equal(x, y)
    assert a.audit() == b.audit()
    a.close()
    b.close()
