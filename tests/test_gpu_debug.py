"""Known-answer entry points of libswe.so, through its C-ABI, against the
oracle's one-step exports, bit for bit: swe_debug_step runs one fly() call
(wind.hpp:28-117) of caller-given particles with the device code of the step
kernels (or_fly_once), swe_debug_cascade one cell's World::cascade
(world.hpp:79-153) with the settling factor of its events (or_cascade_cell),
and swe_set_rng sets the particles' seed (World::SEED).  The particles are
drawn to reach every branch of fly(): fresh and aged, grounded and
airborne, on slopes and over a drop, loaded past the drag's sign change,
past max_age, and off a bounded map."""
import ctypes
import sys

import numpy as np
import pytest

import oracle as O
from simplewinderosion_amd import Batch, World, Wind, capi

pytestmark = pytest.mark.gpu

NX, NY = 24, 20


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def terrain(seed):
    """A tilted rough map with a 0.3 drop across it."""
    rng = np.random.default_rng(seed)
    x = np.arange(NX)[:, None]
    y = np.arange(NY)[None, :]
    h = 0.004 * x + 0.002 * y + 0.05 * rng.random((NX, NY))
    h[NX // 2:, :] += 0.3
    return h.astype(np.float32)


def particles(n, seed):
    rng = np.random.default_rng(seed)
    st = np.zeros((n, 7), np.float32)
    st[:, 0] = rng.uniform(-1.5, NX + 0.5, n)  # some off a bounded map
    st[:, 2] = rng.uniform(-1.5, NY + 0.5, n)
    st[:, 1] = rng.uniform(-0.2, 1.2, n)  # below the ground (clamped) or airborne
    st[:, 3:6] = rng.normal(0.0, 0.6, (n, 3))
    st[::17, 3:6] *= 40.0  # fast particles
    st[:, 6] = rng.choice(np.array([0.0, 0.01, 0.2, 0.9, 3.5, 5.0], np.float32), n)  # 3.5, 5: drag < 0
    st[:, 6] += rng.uniform(0.0, 0.05, n).astype(np.float32) * (st[:, 6] > 0)
    age = rng.choice(np.array([0, 1, 2, 7, 300, Wind.maxAge, Wind.maxAge + 1], np.int32), n)
    key = rng.integers(0, 2 ** 63, n, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, n, dtype=np.uint64)
    return st, age, key


@pytest.mark.parametrize("periodic,layers", [(False, 1), (True, 1), (False, 2), (True, 2)])
def test_debug_step_matches_oracle(periodic, layers):
    say("case debug_step periodic=%d layers=%d" % (periodic, layers))
    h = terrain(3)
    sed = None
    if layers == 2:
        sed = np.random.default_rng(4).uniform(0.0, 0.1, (NX, NY)).astype(np.float32)
        sed[::3, ::2] = 0.0  # bare bedrock: the abrasion branch
    w = World(NX, NY, periodic=periodic, layers=layers, abrasion=0.02 if layers == 2 else 0.0)
    w.upload(h, sed)
    d = w.download()
    S = d["height"]
    SS = d["sediment"] if layers == 2 else None
    op = O.params_from(w.params())
    st, age, key = particles(400, 5 + 2 * layers + periodic)
    g_st, g_age, g_alive, g_pick = w.debug_step(st, age, key)
    stepped = 0
    for i in range(len(st)):
        ok, o_st, o_age, _, o_pick = O.fly_once(op, S, st[i], age[i], int(key[i]), sed_snapshot=SS)
        assert (bool(g_alive[i]), int(g_age[i]), int(g_pick[i])) == (ok, o_age, o_pick), i
        assert np.array_equal(g_st[i].view(np.uint32), o_st.view(np.uint32)), (i, g_st[i], o_st)
        stepped += ok
    # every kind of outcome occurred
    assert 0 < stepped < len(st)
    assert (g_pick > 0).any() and (g_pick < 0).any()


@pytest.mark.parametrize("periodic,layers", [(False, 1), (True, 2)])
def test_debug_cascade_matches_oracle(periodic, layers):
    say("case debug_cascade periodic=%d layers=%d" % (periodic, layers))
    h = terrain(7)
    sed = None
    if layers == 2:
        sed = np.random.default_rng(8).uniform(0.0, 0.05, (NX, NY)).astype(np.float32)
    w = World(NX, NY, periodic=periodic, layers=layers)
    w.upload(h, sed)
    d = w.download()
    S = np.ascontiguousarray(d["height"], np.float32)
    SS = np.ascontiguousarray(d["sediment"], np.float32) if layers == 2 else None
    op = O.params_from(w.params())
    fp = ctypes.POINTER(ctypes.c_float)
    moved = 0
    cells = [(0, 0), (0, NY - 1), (NX - 1, 0), (NX - 1, NY - 1), (5, 0), (0, 7), (NX // 2 - 1, 9), (NX // 2, 9),
             (NX // 2, 0), (3, 4), (17, 13)]
    for (x, y) in cells:
        for events in (1, 2, 7, 40):
            got = w.debug_cascade(x, y, events)
            nbc = np.zeros(8, np.int64)
            nbf = np.zeros(8, np.int64)
            n = O.load().or_cascade_cell(S.ravel().ctypes.data_as(fp),
                                         None if SS is None else SS.ravel().ctypes.data_as(fp),
                                         NX, NY, ctypes.byref(op), x, y, events,
                                         nbc.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                         nbf.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
            want = [(int(nbc[i]), int(nbf[i])) for i in range(n)]
            assert got == want, (x, y, events, got, want)
            moved += sum(1 for _, f in got if f != 0)
    assert moved > 0


def test_set_rng_is_the_seed():
    say("case set_rng")
    h = terrain(9)
    lib = capi.load()
    a = World(NX, NY, periodic=True)
    a.upload(h)
    assert lib.swe_set_rng(a.ctx, 987654321) == 0
    p = capi.SweParams()
    assert lib.swe_get_params(a.ctx, ctypes.byref(p)) == 0 and p.seed == 987654321
    assert lib.swe_erode(a.ctx, 300, 2) == 0  # the C-ABI alone: nothing re-pushes World.SEED
    old = World.SEED
    try:
        World.SEED = 987654321
        b = World(NX, NY, periodic=True)
        b.upload(h)
        b.erode(300, n_cycles=2)
        c = World(NX, NY, periodic=True)
        c.upload(h)
        c.set_rng(5)  # the Python mirror sets the static, which erode pushes
        c.erode(300, n_cycles=2)
    finally:
        World.SEED = old
    ha, hb, hc = a.download()["height"], b.download()["height"], c.download()["height"]
    assert np.array_equal(ha, hb)
    assert not np.array_equal(ha, hc)


def test_debug_entries_refuse_bad_calls():
    """A batch context and a cell outside the map are refused before any
    launch, with the reason in swe_last_error."""
    say("case debug refusals")
    lib = capi.load()
    w = World(NX, NY)
    w.upload(terrain(1))
    cnt = ctypes.c_int32()
    cell = np.zeros(8, np.int64)
    fx = np.zeros(8, np.int64)
    i64 = ctypes.POINTER(ctypes.c_int64)
    for (x, y) in ((-1, 0), (0, NY), (NX, 3)):
        rc = lib.swe_debug_cascade(w.ctx, x, y, 1, ctypes.byref(cnt), cell.ctypes.data_as(i64),
                                   fx.ctypes.data_as(i64))
        assert rc == -1 and b"out of the map" in lib.swe_last_error(w.ctx)
    b = Batch([(32, 32), (48, 40)])
    st, age, key = particles(4, 1)
    i32 = ctypes.POINTER(ctypes.c_int32)
    alive = np.zeros(4, np.int32)
    pick = np.zeros(4, np.int64)
    rc = lib.swe_debug_step(b.ctx, 4, st.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), age.ctypes.data_as(i32),
                            key.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), alive.ctypes.data_as(i32),
                            pick.ctypes.data_as(i64))
    assert rc == -1 and b"map context only" in lib.swe_last_error(b.ctx)
    assert w.debug_step(st[:0], age[:0], key[:0])[2].size == 0  # no particles: nothing to do
