"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_fixtures.py) against the oracle: the batched method must
keep reproducing them bit for bit, and each fixture satisfies the exact
mass ledger on its own."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_fixtures as G  # noqa: E402
import oracle as O  # noqa: E402


def load(name):
    with np.load(os.path.join(HERE, "golden", name + ".npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def exact_sum(a):
    return sum(int(x) for x in np.asarray(a, np.int64).ravel())


@pytest.mark.parametrize("name", sorted(G.CASES))
def test_oracle_reproduces_fixture(name):
    ref = load(name)
    got = G.run(name)
    assert np.array_equal(got["height_in"], ref["height_in"])
    assert np.array_equal(got["zm"], ref["zm"])
    assert np.array_equal(got["ema"], ref["ema"])
    assert int(got["lost"]) == int(ref["lost"]) and int(got["steps"]) == int(ref["steps"])
    if "sm" in ref:
        assert np.array_equal(got["sm"], ref["sm"])


@pytest.mark.parametrize("name", sorted(G.CASES))
def test_fixture_mass_ledger(name):
    ref = load(name)
    m0 = exact_sum(O.fx_array(ref["height_in"].ravel()))
    assert exact_sum(ref["zm"]) + int(ref["lost"]) == m0
    nx, ny, particles, cycles, kw = G.CASES[name]
    if kw.get("periodic"):
        assert int(ref["steps"]) == particles * 513 * cycles
    else:
        assert 0 < int(ref["steps"]) < particles * 513 * cycles
    if kw.get("deposit_on_death"):
        assert int(ref["lost"]) == 0
