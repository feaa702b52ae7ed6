"""GPU parity: libswe.so (through its C-ABI) against the batched oracle, bit for
bit on the int64 master layers, the EMA fields and the carried-off mass."""
import sys

import numpy as np
import pytest

import oracle as O
from simplewinderosion_amd import World, Wind, capi

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def exact_sum(a):
    a = np.asarray(a, np.int64).ravel()
    return int((a >> 32).sum()) * 2 ** 32 + int((a & 0xFFFFFFFF).sum())


def run_pair(nx, ny, particles, cycles, seed=1, sed=None, map_seed=0, height=None, settling=None, path=0,
             **opts):
    """Run the same workload on the GPU and in the oracle; return both states."""
    if settling is not None:
        World.settling = settling
    rng = np.random.default_rng(map_seed)
    h = height if height is not None else (rng.random((nx, ny)) * 0.5).astype(np.float32)
    w = World(nx, ny, **{k: v for k, v in opts.items() if k in
                         ("periodic", "track", "layers", "window", "abrasion", "deposit_on_death",
                          "cascade_order")})
    w.upload(h, sed)
    w.set_path(path)
    World.SEED = seed
    op = O.params_from(w.params())
    st = O.OracleState(nx, ny, h, sediment=sed, params=op)
    st.erode(particles, cycles)
    w.erode(particles, n_cycles=cycles)
    w.synchronize()
    return w, st


def assert_same(w, st):
    zm, sm = w.master()
    assert np.array_equal(zm.ravel(), st.zm), "height masters differ: %d cells" % int((zm.ravel() != st.zm).sum())
    if st.p.layers == 2:
        assert np.array_equal(sm.ravel(), st.sm), "sediment masters differ"
    d = w.download()
    assert np.array_equal(d["height"].ravel(), O.unfx_vec(st.zm))
    if st.p.track:
        n = w.nx * w.ny
        assert np.array_equal(d["massflow"].ravel(), st.ema[:n]), "massflow differs"
        assert np.array_equal(d["momentum"].ravel(), st.ema[n:]), "momentum differs"
    a = w.audit()
    assert a["lost"] == st.lost.value
    assert a["height_sum"] == exact_sum(st.zm)
    assert np.abs(O.unfx_vec(st.zm)).max() < 100.0, "unstable parameters: the map diverged"
    assert a["cycle"] == st.cycle.value


# Cascade gain per window is about 2 * (particles * window / cells) * settling;
# each case keeps it below 0.5 so that the map stays bounded.
CASES = [
    # name, nx, ny, particles, cycles, opts
    ("periodic-ref-params", 32, 32, 20, 2, dict(periodic=True, window=1)),
    ("bounded-ref-params", 64, 64, 40, 2, dict(periodic=False, window=1)),
    ("bounded-w16", 48, 40, 300, 2, dict(periodic=False, window=16, settling=0.05)),
    ("bounded-window1", 20, 24, 100, 1, dict(periodic=False, window=1, settling=0.2)),
    ("periodic-notrack", 40, 33, 300, 2, dict(periodic=True, track=False, window=9, settling=0.05)),
    ("deposit", 32, 32, 300, 2, dict(periodic=True, deposit_on_death=True, window=16, settling=0.02)),
    ("rect-nonpow2", 37, 129, 700, 2, dict(periodic=True, window=5, settling=0.1)),
    ("tall-bounded", 130, 9, 400, 1, dict(periodic=False, window=32, settling=0.01)),
    ("per-cycle-window", 96, 80, 100, 2, dict(periodic=True, window=0, settling=0.01)),
]


@pytest.mark.parametrize("name,nx,ny,particles,cycles,opts", CASES, ids=[c[0] for c in CASES])
def test_parity_single_layer(statics, name, nx, ny, particles, cycles, opts):
    say("case", name, nx, ny, particles, cycles, opts)
    w, st = run_pair(nx, ny, particles, cycles, seed=17, **opts)
    assert_same(w, st)
    w.close()


def test_parity_two_layer_abrasion(statics):
    say("case two-layer 64x64")
    Wind.suspension = 1e-2
    World.settling = 0.1
    sed = np.full((64, 64), 0.003, np.float32)
    w, st = run_pair(64, 64, 800, 3, seed=5, sed=sed, periodic=True, layers=2, abrasion=1e-3, window=8, settling=0.1)
    assert_same(w, st)
    zm, sm = w.master()
    assert sm.min() >= 0 and (sm == 0).any()
    w.close()


@pytest.mark.parametrize("window", [1, 4])
@pytest.mark.parametrize("nx,ny,periodic", [(1, 1, True), (1, 1, False), (2, 3, True), (3, 2, False), (1, 70, True),
                                            (20, 40, True), (33, 65, True), (17, 33, False)])
def test_parity_small_maps(statics, nx, ny, periodic, window):
    """Tiny maps and maps of one to three 16x32 tiles per axis (tile-direction
    wrap-around on the tile-owned path)."""
    say("case small", nx, ny, periodic, window)
    w, st = run_pair(nx, ny, 3 * nx * ny // 4 + 5, 2, seed=3, periodic=periodic, window=window, settling=0.01)
    assert_same(w, st)
    w.close()


@pytest.mark.parametrize("periodic,layers,track,deposit", [(True, 1, True, False), (False, 1, True, True),
                                                           (True, 2, False, False), (False, 2, True, True)])
def test_both_paths_agree_at_window_one(statics, periodic, layers, track, deposit):
    """The tile-owned step kernel and the global-memory march give the same
    bits, and both equal the oracle."""
    say("case paths", periodic, layers, track, deposit)
    Wind.suspension = 1e-2
    sed = np.full((70, 100), 0.002, np.float32) if layers == 2 else None
    outs = []
    for path in (1, 2):
        w, st = run_pair(70, 100, 4000, 2, seed=9, sed=sed, periodic=periodic, layers=layers, track=track,
                         deposit_on_death=deposit, abrasion=1e-3, window=1, settling=0.05, path=path)
        assert_same(w, st)
        outs.append(w.master())
        w.close()
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("order", [1, 3, 64])
def test_tile_order_does_not_change_results(statics, order):
    """The workgroup -> tile order is a speed knob: every order gives the
    oracle's bits (170x230 has 11x8 = 88 tiles, so orders 3 and 64 also
    exercise the tail that keeps its own index)."""
    from simplewinderosion_amd import capi

    say("case tile order", order)
    Wind.suspension = 1e-2
    World.settling = 0.05
    rng = np.random.default_rng(5)
    h = (rng.random((170, 230)) * 0.5).astype(np.float32)
    w = World(170, 230, periodic=True, layers=2, window=1, abrasion=1e-3)
    sed = np.full((170, 230), 0.002, np.float32)
    w.upload(h, sed)
    w.set_tune(capi.SWE_TUNE_TILE_ORDER, order)
    assert w.get_tune(capi.SWE_TUNE_TILE_ORDER) == order
    World.SEED = 3
    st = O.OracleState(170, 230, h, sediment=sed, params=O.params_from(w.params()))
    st.erode(30000, 2)
    w.erode(30000, n_cycles=2)
    w.synchronize()
    assert_same(w, st)
    w.close()


@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("path", [1, 2])
def test_step_statistics(statics, periodic, path):
    """swe_stats counts the fly() calls that returned true on the device: the
    oracle's count on bounded maps (early deaths) and 513 per particle per
    cycle on periodic ones; the tile statistics are consistent."""
    say("case stats", periodic, path)
    Wind.suspension = 1e-2
    w, st = run_pair(70, 100, 4000, 2, seed=9, periodic=periodic, window=1, settling=0.05, path=path)
    assert_same(w, st)
    s = w.stats()
    assert s["particle_steps"] == st.steps.value
    if periodic:
        assert s["particle_steps"] == 4000 * 513 * 2
    if path == 2:
        assert 0 < s["max_tile_particles"] <= 4000
        assert s["crowded_tile_windows"] == 0 or s["max_tile_particles"] > 2048
    w.reset_stats()
    s = w.stats()
    assert s["particle_steps"] == 0 and s["max_tile_particles"] == 0
    w.close()


@pytest.mark.parametrize("nx,ny,particles,settling", [(64, 64, 20000, 0.02), (32, 32, 12000, 0.01)])
def test_crowded_tiles(statics, nx, ny, particles, settling):
    """Tiles holding more particles than the step kernel's LDS list (2048)
    take their sources in several lists; results stay the oracle's bits.
    32x32 has two tiles of ~6000 particles: six lists per tile and window."""
    say("case crowded", nx, ny, particles)
    Wind.suspension = 1e-4
    w, st = run_pair(nx, ny, particles, 1, seed=4, periodic=True, window=1, settling=settling, path=2,
                     layers=2, abrasion=1e-3, sed=np.full((nx, ny), 0.002, np.float32))
    assert_same(w, st)
    s = w.stats()
    assert s["crowded_tile_windows"] > 0 and s["max_tile_particles"] > 2048
    assert s["particle_steps"] == st.steps.value
    w.close()


@pytest.mark.parametrize("particles,modes,periodic,order", [
    (90000, (1, 0, 2), True, 0),
    (20000, (2,), True, 0),
    (20000, (2,), False, 0),  # bounded map: particles leave, tiles at the edges
    (20000, (2,), True, 4),   # light tiles' home XCD under a round-robin tile order
])
def test_heavy_first_tiles(statics, particles, modes, periodic, order):
    """SWE_TUNE_HEAVY_FIRST: tiles that received more than 1024 particles in a
    window run on the first workgroups of the next launch (mode 1), and in
    mode 2, the default up to 4096 tiles, tiles with at most 256 also run
    on the last ones, listed per XCD.  128x256 is 64 tiles: 8 heavy and 32 light slots, so tiles
    are listed, skipped and left over (~1400 particles per tile at 90000,
    ~300 at 20000).  Results are the oracle's bits in every mode."""
    say("case heavy-first", particles, modes, periodic, order)
    Wind.suspension = 1e-4
    World.settling = 0.02
    World.SEED = 6
    nx, ny = 128, 256
    sed = np.full((nx, ny), 0.002, np.float32)
    h = (np.random.default_rng(0).random((nx, ny)) * 0.5).astype(np.float32)
    st = None
    for mode in modes:
        w = World(nx, ny, periodic=periodic, window=1, layers=2, abrasion=1e-3)
        w.upload(h, sed)
        w.set_path(2)
        w.set_tune(capi.SWE_TUNE_TILE_ORDER, order)
        if st is None:
            st = O.OracleState(nx, ny, h, sediment=sed, params=O.params_from(w.params()))
            st.erode(particles, 1)
        assert w.get_tune(capi.SWE_TUNE_HEAVY_FIRST) == 2  # the default at 64 tiles
        w.set_tune(capi.SWE_TUNE_HEAVY_FIRST, mode)
        assert w.get_tune(capi.SWE_TUNE_HEAVY_FIRST) == mode
        w.erode(particles, n_cycles=1)
        w.synchronize()
        assert_same(w, st)
        s = w.stats()
        assert s["particle_steps"] == st.steps.value
        if mode == 0:
            assert s["heavy_tile_windows"] == 0 and s["light_tile_windows"] == 0
        if mode >= 1 and particles == 90000:
            assert s["heavy_tile_windows"] > 0 and s["max_tile_particles"] > 1024
        if mode == 1:
            assert s["light_tile_windows"] == 0
        if mode == 2 and particles == 20000:
            assert s["light_tile_windows"] > 0
        w.close()


def test_parity_cpu256_workload(statics):
    """The 256x256 two-layer workload (fBm + sediment 0.05, wind (1,0,1)) for 3 cycles."""
    from simplewinderosion_amd import workloads

    say("case cpu256 3 cycles")
    w = workloads.make_world("cpu256", seed=4)
    h, s = workloads.make_inputs("cpu256", seed=4)
    st = O.OracleState(256, 256, h, sediment=s, params=O.params_from(w.params()))
    st.erode(2000, 3)
    w.erode(2000, n_cycles=3)
    w.synchronize()
    assert_same(w, st)
    w.close()


def test_repeatable_and_split_calls_equal_one_call(statics):
    say("case split calls")
    rng = np.random.default_rng(1)
    h = (rng.random((50, 50)) * 0.5).astype(np.float32)
    outs = []
    World.settling = 0.02
    for split in (False, True, False):
        w = World(50, 50, periodic=True, window=16)
        w.upload(h)
        if split:
            w.erode(400)
            w.erode(400)
        else:
            w.erode(400, n_cycles=2)
        outs.append(w.master()[0])
        w.close()
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_normals_match_oracle(statics):
    import ctypes
    say("case normals")
    rng = np.random.default_rng(2)
    for periodic in (False, True):
        h = (rng.random((17, 23)) * 0.2).astype(np.float32)
        w = World(17, 23, periodic=periodic)
        w.upload(h)
        n = w.normals()
        S = O.unfx_vec(O._fx_vec(h)).reshape(17, 23)
        out = np.zeros(3, np.float32)
        for x in range(17):
            for y in range(23):
                O.load().or_normal(S.ravel().ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 17, 23, int(periodic),
                                   80.0, x, y, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
                assert np.array_equal(n[:, x, y], out), (x, y)
        w.close()


def test_full_size_one_cycle(statics):
    """1024x1024 two-layer map with pyramids, 1,000,000 particles, one cycle,
    one-call snapshot windows (the benchmark workload): bit-exact against the
    oracle on all host cores."""
    from simplewinderosion_amd import workloads

    say("case gpu1024 1 cycle")
    w = workloads.make_world("gpu1024", seed=1)
    h, s = workloads.make_inputs("gpu1024", seed=1)
    w.erode(1_000_000, n_cycles=1)
    w.synchronize()
    st = O.OracleState(1024, 1024, h, sediment=s, params=O.params_from(w.params()))
    st.erode(1_000_000, 1)
    assert_same(w, st)
    w.close()


def test_checkpoint_resume_is_bit_exact(statics, tmp_path):
    """save after 2 cycles, continue 2 more; load the checkpoint into a fresh
    context and run the same 2: identical masters, fields and counters."""
    say("case checkpoint")
    Wind.suspension = 1e-2
    World.settling = 0.05
    World.SEED = 21
    rng = np.random.default_rng(4)
    h = (rng.random((40, 72)) * 0.5).astype(np.float32)
    sed = np.full((40, 72), 0.003, np.float32)
    w = World(40, 72, periodic=True, layers=2, abrasion=1e-3)
    w.upload(h, sed)
    w.erode(900, n_cycles=2)
    path = tmp_path / "state.swe"
    w.save(path)
    w.erode(900, n_cycles=2)
    ref = (w.master(), w.download(), w.audit())
    w.close()
    World.settling = 0.7  # the checkpoint's params must win
    w2 = World(40, 72)
    w2.load(path)
    w2.erode(900, n_cycles=2)
    got = (w2.master(), w2.download(), w2.audit())
    assert np.array_equal(ref[0][0], got[0][0]) and np.array_equal(ref[0][1], got[0][1])
    for k in ("massflow", "momentum", "height", "sediment"):
        assert np.array_equal(ref[1][k], got[1][k]), k
    assert ref[2] == got[2]
    # and the resumed run is the oracle's four cycles
    st = O.OracleState(40, 72, h, sediment=sed, params=O.params_from(w2.params()))
    st.erode(900, 4)
    assert_same(w2, st)
    w2.close()


@pytest.mark.parametrize("periodic,nx,ny", [(True, 96, 128), (False, 112, 100), (True, 40, 200)])
def test_far_particle_routing(statics, periodic, nx, ny):
    """A fast prevailing wind moves particles more than one tile (16 x 32
    cells) per call: they travel through the far list that every workgroup
    scans.  Results stay the oracle's bits, and the device counts far moves."""
    say("case far", periodic, nx, ny)
    Wind.suspension = 1e-3
    Wind.pspeed = (40.0, 0.0, 35.0)
    w, st = run_pair(nx, ny, 3000, 2, seed=13, periodic=periodic, window=1, settling=0.05, path=2)
    assert_same(w, st)
    s = w.stats()
    assert s["far_particles"] > 0
    assert s["particle_steps"] == st.steps.value
    w.close()


@pytest.mark.parametrize("window", [1, 3])
def test_parity_edge_spawn(statics, window):
    """The commented boundary spawn of world.hpp:56-61 (spawn = 1)."""
    say("case edge spawn", window)
    rng = np.random.default_rng(8)
    h = (rng.random((60, 50)) * 0.3).astype(np.float32)
    World.settling = 0.05
    World.SEED = 4
    w = World(60, 50, periodic=False, window=window, spawn=1)
    w.upload(h)
    st = O.OracleState(60, 50, h, params=O.params_from(w.params()))
    st.erode(400, 2)
    w.erode(400, n_cycles=2)
    assert_same(w, st)
    w.close()


def test_windmap_matches_erf(statics):
    """SimpleWindErosion.cpp:224-230 colour of the EMA momentum field."""
    import math
    say("case windmap")
    rng = np.random.default_rng(3)
    w = World(21, 34)
    mom = (rng.standard_normal(3 * 21 * 34) * 0.7).astype(np.float32)
    w.upload(np.zeros((21, 34), np.float32))
    w.upload_fields(momentum=mom)
    got = w.windmap()
    m = mom.reshape(3, 21, 34)
    for axis in range(3):
        want = np.vectorize(lambda v: 0.5 * (1.0 + math.erf(2.0 * float(np.float32(v)))))(m[axis])
        assert np.abs(got[..., axis] - want).max() < 3e-7
    assert np.all(got[..., 3] == 1.0)
    w.close()


@pytest.mark.parametrize("periodic", [False, True])
def test_mesh_attributes(statics, periodic):
    """updatenode (cellpool.h:302-333): P, normal, T - P, B - P per cell."""
    import ctypes
    say("case mesh", periodic)
    rng = np.random.default_rng(6)
    nx, ny = 13, 19
    h = (rng.random((nx, ny)) * 0.2).astype(np.float32)
    w = World(nx, ny, periodic=periodic)
    w.upload(h)
    m = w.mesh()
    S = O.unfx_vec(O._fx_vec(h)).reshape(nx, ny)
    F = np.float32
    out = np.zeros(3, np.float32)
    for x in range(nx):
        for y in range(ny):
            hp = S[x, y]
            hx = S[(x + 1) % nx, y] if (periodic or x + 1 < nx) else F(0)
            hy = S[x, (y + 1) % ny] if (periodic or y + 1 < ny) else F(0)
            P = np.array([x, F(80) * hp, y], np.float32)
            T = np.array([x + 1, F(80) * hx, y], np.float32) - P
            B = np.array([x, F(80) * hy, y + 1], np.float32) - P
            O.load().or_normal(S.ravel().ctypes.data_as(ctypes.POINTER(ctypes.c_float)), nx, ny, int(periodic),
                               80.0, x, y, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
            assert np.array_equal(m[x, y], np.concatenate([P, out, T, B])), (x, y)
    w.close()
