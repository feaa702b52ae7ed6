"""The colour-ordered cascade on the GPU (cascade_order 1, or automatic above
settling 0.25): bit-exact against the oracle, including the reference's own
defaults at its density (4096 particles on 512^2, SimpleWindErosion.cpp:217)."""
import numpy as np
import pytest

import oracle as O
from simplewinderosion_amd import Batch, World, Wind, glm_normalize
from simplewinderosion_amd.world import SweError

from test_gpu_parity import assert_same, run_pair, say

pytestmark = pytest.mark.gpu

CASES = [
    # name, nx, ny, particles, cycles, options
    ("bounded-auto-s1", 40, 37, 300, 2, dict(periodic=False, window=1, settling=1.0)),
    ("periodic-2layer-s1", 32, 29, 500, 2, dict(periodic=True, window=1, settling=1.0, layers=2, abrasion=1e-3)),
    ("periodic-tiny-1x7", 1, 7, 20, 2, dict(periodic=True, window=1, settling=1.0)),
    ("periodic-tiny-2x5", 2, 5, 20, 2, dict(periodic=True, window=1, settling=1.0)),
    ("bounded-window4", 48, 30, 400, 2, dict(periodic=False, window=4, settling=0.6, cascade_order=1)),
    ("periodic-explicit-low", 64, 64, 2000, 1, dict(periodic=True, window=1, settling=0.05, cascade_order=1,
                                                     deposit_on_death=True)),
]


@pytest.mark.parametrize("path", [1, 2])
@pytest.mark.parametrize("name,nx,ny,particles,cycles,opts", CASES, ids=[c[0] for c in CASES])
def test_colour_cascade_matches_oracle(statics, name, nx, ny, particles, cycles, opts, path):
    """Both execution paths: the global march with per-class launches, and
    the tile step kernel moving the particles (window 1) with the same
    per-class launches after each window."""
    if path == 2 and opts.get("window", 1) != 1:
        pytest.skip("the tile step path runs window 1")
    say("case colour", name, "path", path)
    Wind.suspension = 0.01
    w, st = run_pair(nx, ny, particles, cycles, seed=7, path=path, **opts)
    assert O.load().or_colour_mode(st.p) == 1
    assert_same(w, st)
    w.close()


def test_colour_tile_path_crowded_and_mode_switch(statics):
    """The benchmark's density (one particle per cell, crowded tiles) on the
    tile path with the colour-ordered cascade, then synchronous cycles on
    the same context: the oracle's bits throughout."""
    say("case colour crowded + switch")
    Wind.suspension = 1e-3
    World.settling = 0.01
    nx, ny = 96, 160
    rng = np.random.default_rng(12)
    h = (rng.random((nx, ny)) * 0.5).astype(np.float32)
    sed = np.full((nx, ny), 0.01, np.float32)
    w, st = run_pair(nx, ny, nx * ny, 2, seed=3, sed=sed, periodic=True, layers=2, abrasion=1e-3, window=1,
                     cascade_order=1, path=2)
    assert_same(w, st)
    # two more cycles with the synchronous cascade on both sides
    w.cascade_order = 0
    st.p.cascade_order = 0
    w.erode(nx * ny, n_cycles=2)
    w.synchronize()
    st.erode(nx * ny, 2)
    assert_same(w, st)
    w.close()


def test_reference_defaults_at_reference_density(statics):
    """Every parameter at the reference's value (bounded 512^2 map, settling
    1, suspension 0.05, maxdiff 0.005, pspeed (1,0,0)), 4096 particles per
    cycle as SimpleWindErosion.cpp:217, two cycles: the oracle's bits, the
    map bounded."""
    say("case reference defaults 512^2 x 4096")
    from simplewinderosion_amd import world as Wm

    h = (Wm.generate_terrain(512, 512, 3, 6, 4.0) * 0.5).astype(np.float32)
    w, st = run_pair(512, 512, 4096, 2, seed=1, height=h)
    assert w.params().settling == 1.0 and w.params().cascade_order == -1
    assert_same(w, st)
    z = O.unfx_vec(st.zm)
    assert z.min() >= h.min() - 0.05 and z.max() <= h.max() + 0.05
    w.close()


def test_synchronous_cascade_refuses_large_settling(statics):
    say("case colour refusals")
    World.settling = 1.0
    with pytest.raises(SweError):
        World(16, 16, cascade_order=0)
    with pytest.raises(SweError):
        Batch([(32, 32), (16, 16)], periodic=True, cascade_order=0)


@pytest.mark.parametrize("periodic", [True, False])
def test_colour_cascade_batch_equals_single_runs(statics, periodic):
    """A batch at the reference's settling (1.0, automatic colour order):
    each class runs on every terrain in one launch.  Every terrain equals
    the same terrain eroded alone (sides of every residue mod 3, so the
    terrains have different class grids, and one 2-cell-wide terrain whose
    neighbourhoods alias), and one terrain equals the oracle."""
    sizes = [(31, 29), (40, 17), (33, 64), (2, 7), (10, 10)]
    say("case colour batch", periodic, sizes)
    Wind.suspension = 1e-3
    World.settling = 1.0
    layers = 2 if periodic else 1
    rng = np.random.default_rng(5)
    winds = [glm_normalize((np.cos(a), 0.2, np.sin(a))) for a in rng.uniform(0, 6.28, len(sizes))]
    seeds = [101, 202, 303, 404, 505]
    heights = [(rng.random(sz) * 0.4).astype(np.float32) for sz in sizes]
    seds = [np.full(sz, 0.004, np.float32) for sz in sizes]
    parts = [int(0.8 * a * b) for a, b in sizes]
    b = Batch(sizes, periodic=periodic, layers=layers, abrasion=1e-3 if layers == 2 else 0.0)
    assert b.params().cascade_order == -1
    for k in range(len(sizes)):
        b.set_terrain(k, seeds[k], winds[k])
        b.upload(k, heights[k], seds[k] if layers == 2 else None)
    b.erode(parts, n_cycles=2)
    for k, (nx, ny) in enumerate(sizes):
        World.SEED = seeds[k]
        Wind.pspeed = winds[k]
        w = World(nx, ny, periodic=periodic, layers=layers, abrasion=1e-3 if layers == 2 else 0.0)
        w.upload(heights[k], seds[k] if layers == 2 else None)
        w.erode(parts[k], n_cycles=2)
        zb, sb = b.master(k)
        zw, sw = w.master()
        assert np.array_equal(zb, zw), k
        if layers == 2:
            assert np.array_equal(sb, sw), k
        assert b.audit(k)["lost"] == w.audit()["lost"], k
        if k == 0:
            st = O.OracleState(nx, ny, heights[k], sediment=seds[k] if layers == 2 else None,
                               params=O.params_from(w.params()))
            assert O.load().or_colour_mode(st.p) == 1
            st.erode(parts[k], 2)
            assert np.array_equal(zw.ravel(), st.zm)
        w.close()
    b.close()
