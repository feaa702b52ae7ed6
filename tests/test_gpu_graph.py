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
