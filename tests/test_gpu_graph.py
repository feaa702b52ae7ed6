"""The tile path's HIP-graph mode (SWE_TUNE_GRAPH = 1, the default) against
individual launches: every replay of the captured cycle, with the program's
own work and synchronisation between replays, gives the eager run's bits and
the oracle's; changing the particle count recaptures."""
import numpy as np
import pytest

import oracle as O
from simplewinderosion_amd import World, Wind, capi

pytestmark = pytest.mark.gpu


def make(graph, h, s):
    w = World(96, 160, periodic=True, layers=2, window=1, abrasion=1e-3)
    w.upload(h, s)
    w.set_tune(capi.SWE_TUNE_GRAPH, graph)
    assert w.get_tune(capi.SWE_TUNE_GRAPH) == graph
    return w


def test_graph_replays_equal_eager_launches(statics):
    print("case graph replays", flush=True)
    Wind.suspension = 1e-3
    World.settling = 0.02
    World.SEED = 21
    rng = np.random.default_rng(3)
    h = (rng.random((96, 160)) * 0.5).astype(np.float32)
    s = np.full((96, 160), 0.01, np.float32)
    g, e = make(1, h, s), make(0, h, s)
    st = O.OracleState(96, 160, h, sediment=s, params=O.params_from(g.params()))
    # five cycles: the two parities are captured, then replayed; a different
    # particle count in the middle forces a new capture
    for n in (6000, 6000, 6000, 3000, 6000):
        g.erode(n)
        e.erode(n)
        st.erode(n, 1)
        zg, sg = g.master()  # synchronises; host work between replays
        ze, se = e.master()
        g.normals()
        assert np.array_equal(zg, ze) and np.array_equal(sg, se)
        assert np.array_equal(zg.ravel(), st.zm) and np.array_equal(sg.ravel(), st.sm)
        assert np.array_equal(g.download()["momentum"], e.download()["momentum"])
    assert g.audit() == e.audit()
    assert g.stats()["particle_steps"] == e.stats()["particle_steps"] == st.steps.value
    g.close()
    e.close()


def test_graph_after_buffers_grow_and_shrink(statics):
    """A larger count reallocates the particle buffers; a later smaller count
    must not replay a graph captured against the freed ones (the allocation
    generation is part of the graph key).  Every cycle is checked against the
    oracle."""
    print("case graph grow/shrink", flush=True)
    Wind.suspension = 1e-3
    World.settling = 0.02
    World.SEED = 5
    rng = np.random.default_rng(8)
    h = (rng.random((96, 160)) * 0.5).astype(np.float32)
    s = np.full((96, 160), 0.01, np.float32)
    g = make(1, h, s)
    st = O.OracleState(96, 160, h, sediment=s, params=O.params_from(g.params()))
    for n in (2000, 2000, 7000, 2000, 2000, 9000, 2000):
        g.erode(n)
        st.erode(n, 1)
        zg, sg = g.master()
        assert np.array_equal(zg.ravel(), st.zm) and np.array_equal(sg.ravel(), st.sm), n
    assert g.stats()["particle_steps"] == st.steps.value
    g.close()


def test_upload_size_is_checked():
    w = World(16, 24, periodic=True)
    with pytest.raises(ValueError):
        w.upload_master(np.zeros(16 * 24 - 1, np.int64))
    with pytest.raises(ValueError):
        w.upload_fields(momentum=np.zeros(16 * 24, np.float32))
    w.upload_fields(massflow=np.zeros((16, 24), np.float32), momentum=np.zeros((3, 16, 24), np.float32))
    w.close()
