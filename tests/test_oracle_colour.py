"""The oracle's colour-ordered cascade (cascade_order 1, the automatic choice
above settling 0.25) against the reference's sequential World::cascade
(world.hpp:79-153) and at the reference's own defaults and density."""
import numpy as np
import pytest

import oracle as O
from simplewinderosion_amd import world as W


@pytest.mark.parametrize("periodic", [0, 1])
def test_colour_classes_partition_the_axis_and_keep_three_apart(periodic):
    for n in range(1, 41):
        classes = O.colour_cells(n, periodic)
        flat = sorted(x for c in classes for x in c)
        assert flat == list(range(n)), n
        for c in classes:
            for i, a in enumerate(c):
                for b in c[i + 1:]:
                    d = abs(a - b)
                    if periodic:
                        d = min(d, n - d)
                    assert d >= 3, (n, c)


def _seq_in_order(p, h0, V, order, sed=None):
    H = h0.copy()
    S = None if sed is None else sed.copy()
    for x, y in order:
        if V[x, y]:
            H, S = O.cascade_seq(p, H, x, y, sediment=S)
    return H, S


def _class_order(nx, ny, periodic):
    cx, cy = O.colour_cells(nx, periodic), O.colour_cells(ny, periodic)
    return [(x, y) for xs in cx for ys in cy for x in xs for y in ys]


@pytest.mark.parametrize("periodic,nx,ny", [(0, 20, 23), (1, 20, 23), (1, 31, 17), (0, 9, 10), (1, 4, 5)])
def test_colour_relax_is_the_sequential_cascades_in_class_order(periodic, nx, ny):
    """One event at half the cells, settling 1: the colour relax equals the
    reference's cascade called cell after cell in class order on a float
    map, to float rounding (the relax sums in fixed point); the same calls
    in row-major order end elsewhere, so the order is what is pinned."""
    rng = np.random.default_rng(nx * 100 + ny + periodic)
    h = (rng.random((nx, ny)) * 0.3).astype(np.float32)
    zm = O._fx_vec(h).reshape(nx, ny)
    h0 = O.unfx_vec(zm.ravel()).reshape(nx, ny).astype(np.float32)
    V = (rng.random((nx, ny)) < 0.5).astype(np.uint32)
    p = O.reference_params(periodic=periodic, settling=1.0)
    assert O.load().or_colour_mode(p) == 1  # automatic above 0.25
    z1, _ = O.relax_colour(p, zm, V)
    got = O.unfx_vec(z1.ravel()).reshape(nx, ny)
    want, _ = _seq_in_order(p, h0, V, _class_order(nx, ny, periodic))
    assert np.abs(got - want).max() <= 1e-6
    assert np.abs(want - h0).max() > 0.05
    if nx >= 9:
        rowmajor, _ = _seq_in_order(p, h0, V, [(x, y) for x in range(nx) for y in range(ny)])
        assert np.abs(rowmajor - want).max() > 1e-3
    # exact fixed-point conservation
    assert int(z1.sum()) == int(zm.sum())


def test_colour_relax_two_layers_caps_by_current_sediment():
    """Two layers: transfers capped by the giver's loose sediment as the
    earlier classes left it (sequential cascade with the sediment map)."""
    nx, ny, periodic = 18, 21, 1
    rng = np.random.default_rng(7)
    h = (rng.random((nx, ny)) * 0.2).astype(np.float32)
    sed = (rng.random((nx, ny)) * 0.01).astype(np.float32)
    zm, sm = O._fx_vec(h).reshape(nx, ny), O._fx_vec(sed).reshape(nx, ny)
    h0 = O.unfx_vec(zm.ravel()).reshape(nx, ny).astype(np.float32)
    s0 = O.unfx_vec(sm.ravel()).reshape(nx, ny).astype(np.float32)
    V = (rng.random((nx, ny)) < 0.6).astype(np.uint32)
    p = O.reference_params(periodic=periodic, settling=1.0, layers=2)
    z1, s1 = O.relax_colour(p, zm, V, sm=sm)
    want_h, want_s = _seq_in_order(p, h0, V, _class_order(nx, ny, periodic), sed=s0)
    assert np.abs(O.unfx_vec(z1.ravel()).reshape(nx, ny) - want_h).max() <= 1e-6
    assert np.abs(O.unfx_vec(s1.ravel()).reshape(nx, ny) - want_s).max() <= 1e-6
    # the surface moves exactly; the sediment layer is clamped at zero
    assert (s1 >= 0).all() and int(z1.sum()) == int(zm.sum())


def test_reference_defaults_at_reference_density_follow_the_sequential_run():
    """The reference's own parameters (settling 1, suspension 0.05, bounded
    map) at its density (4096 particles on 512^2, SimpleWindErosion.cpp:217:
    here 144 per cycle on 96^2) for 20 cycles: the batched method with the
    automatic (colour-ordered) cascade stays bounded and ends as close to
    the sequential run as two sequential runs of different seeds are to
    each other."""
    n, cycles, npart = 96, 20, 96 * 96 // 64
    h = (W.generate_terrain(n, n, 7, 6, 4.0) * 0.5).astype(np.float32)

    def seq(seed):
        hs = h.copy().ravel()
        O.load().or_srand(seed)
        O.erode_seq(O.reference_params(periodic=0), n, n, hs, particles=npart, n_cycles=cycles)
        return hs.reshape(n, n)

    p = O.reference_params(periodic=0, seed=1)
    assert p.settling == 1.0 and abs(p.suspension - 0.05) < 1e-9 and p.cascade_order == -1
    st = O.OracleState(n, n, h, params=p)
    st.erode(npart, cycles, nthreads=1)  # 144 particles: threads would only wait on each other
    got = O.unfx_vec(st.zm).reshape(n, n)
    s1, s2 = seq(1), seq(2)
    rms = lambda a, b: float(np.sqrt(np.mean((a - b) ** 2)))  # noqa: E731
    spread = rms(s1, s2)
    assert rms(got, s1) <= 1.5 * spread and rms(got, s2) <= 1.5 * spread, (rms(got, s1), rms(got, s2), spread)
    assert got.min() >= h.min() - 0.01 and got.max() <= h.max() + 0.01
    assert 0.8 <= rms(got, h) / rms(s1, h) <= 1.25
