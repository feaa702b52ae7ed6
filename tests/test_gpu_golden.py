"""libswe.so (through its C-ABI) against the committed golden fixtures of
the batched method (tests/golden/*.npz): same seeded inputs, same bits."""
import os
import sys

import numpy as np
import pytest

from simplewinderosion_amd import World, Wind

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_fixtures as G  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", sorted(G.CASES))
def test_gpu_matches_fixture(statics, name):
    print("case golden", name, flush=True)
    with np.load(os.path.join(HERE, "golden", name + ".npz"), allow_pickle=False) as f:
        ref = {k: f[k] for k in f.files}
    nx, ny, particles, cycles, kw = G.CASES[name]
    Wind.suspension = kw["suspension"]
    World.settling = kw["settling"]
    World.SEED = kw["seed"]
    w = World(nx, ny, periodic=bool(kw.get("periodic")), layers=kw.get("layers", 1), window=kw["window"],
              abrasion=kw.get("abrasion", 0.0), deposit_on_death=bool(kw.get("deposit_on_death")),
              spawn=kw.get("spawn", 0))
    w.upload(ref["height_in"], ref.get("sediment_in"))
    w.erode(particles, n_cycles=cycles)
    w.synchronize()
    zm, sm = w.master()
    assert np.array_equal(zm.ravel(), ref["zm"])
    if "sm" in ref:
        assert np.array_equal(sm.ravel(), ref["sm"])
    d = w.download()
    n = nx * ny
    assert np.array_equal(d["massflow"].ravel(), ref["ema"][:n])
    assert np.array_equal(d["momentum"].ravel(), ref["ema"][n:])
    a = w.audit()
    assert a["lost"] == int(ref["lost"]) and a["cycle"] == int(ref["cycle"])
    assert w.stats()["particle_steps"] == int(ref["steps"])
    w.close()
