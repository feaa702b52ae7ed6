"""Pins of the oracle's surface normal (cellpool.h:200-223) and cascade event
(world.hpp:79-153) by closed forms and hand-computed cases."""
import ctypes

import numpy as np
import pytest

import oracle as O

F32 = np.float32


def normal(S, periodic, x, y, mapscale=80.0):
    S = np.ascontiguousarray(S, np.float32)
    out = np.zeros(3, np.float32)
    O.load().or_normal(S.ravel().ctypes.data_as(ctypes.POINTER(ctypes.c_float)), S.shape[0], S.shape[1],
                       int(periodic), float(mapscale), int(x), int(y),
                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return out


def test_flat_map_normal_is_exactly_up():
    S = np.full((8, 9), 0.37, np.float32)
    for periodic in (0, 1):
        for x, y in [(0, 0), (3, 4), (7, 8), (0, 8), (7, 0)]:
            assert list(normal(S, periodic, x, y)) == [0.0, 1.0, 0.0]


@pytest.mark.parametrize("a", [1.0 / 128, -1.0 / 64, 3.0 / 256])
def test_ramp_in_x(a):
    # h = a*x -> every plane gives (-mapscale*a, 1, 0)
    nx, ny = 16, 8
    S = (np.arange(nx, dtype=np.float32)[:, None] * np.float32(a) + np.zeros((1, ny), np.float32)).astype(np.float32)
    want = np.array([-80.0 * a, 1.0, 0.0])
    want /= np.linalg.norm(want)
    n = normal(S, 0, 5, 3)
    assert np.allclose(n, want, rtol=0, atol=3e-7)
    assert n[2] == 0.0


@pytest.mark.parametrize("b", [1.0 / 128, -1.0 / 32])
def test_ramp_in_y(b):
    nx, ny = 8, 16
    S = (np.zeros((nx, 1), np.float32) + np.arange(ny, dtype=np.float32)[None, :] * np.float32(b)).astype(np.float32)
    want = np.array([0.0, 1.0, -80.0 * b])
    want /= np.linalg.norm(want)
    n = normal(S, 0, 4, 7)
    assert np.allclose(n, want, rtol=0, atol=3e-7)
    assert n[0] == 0.0


def test_bounded_corner_uses_one_plane():
    # at (0,0) only the plane gated by (1,1) exists: cross(s*(0,hN-h,1), s*(1,hE-h,0))
    S = np.zeros((4, 4), np.float32)
    S[0, 1] = 0.01   # north  (y+1)
    S[1, 0] = 0.02   # east   (x+1)
    n = normal(S, 0, 0, 0)
    A, B = 80 * 0.01, 80 * 0.02
    want = np.array([-B, 1.0, -A])
    want /= np.linalg.norm(want)
    assert np.allclose(n, want, atol=3e-7)
    # the same cell on a periodic map also sees the wrapped west/south neighbours
    S[3, 0] = 0.05
    assert not np.allclose(normal(S, 1, 0, 0), n, atol=1e-4)
    assert np.allclose(normal(S, 0, 0, 0), n, atol=0)


def test_periodic_wrap_equals_shifted_interior():
    rng = np.random.default_rng(5)
    S = (rng.random((6, 7)) * 0.1).astype(np.float32)
    for x in range(6):
        for y in range(7):
            big = np.tile(S, (3, 3))
            a = normal(S, 1, x, y)
            b = normal(big, 0, x + 6, y + 7)
            assert np.array_equal(a, b)


# -- cascade -----------------------------------------------------------------

def cascade(S, x, y, SS=None, events=1, **kw):
    p = O.reference_params(**kw)
    S = np.ascontiguousarray(S, np.float32)
    nbc = np.zeros(8, np.int64)
    nbf = np.zeros(8, np.int64)
    ss = None if SS is None else np.ascontiguousarray(SS, np.float32)
    n = O.load().or_cascade_cell(S.ravel().ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                 None if ss is None else ss.ravel().ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                 S.shape[0], S.shape[1], ctypes.byref(p), int(x), int(y), int(events),
                                 nbc.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                 nbf.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    return [(int(nbc[i]), int(nbf[i])) for i in range(n) if nbf[i] != 0]


def test_pair_transfer():
    # 1x2 bounded map: the only neighbour of (0,0) is (0,1), distance 1
    S = np.array([[0.1, 0.0]], np.float32)
    got = cascade(S, 0, 0)
    excess = F32(F32(0.1) - F32(0.0)) - F32(F32(1.0) * F32(0.005))
    t = F32(F32(1.0) * excess) / F32(2.0)
    assert got == [(1, O.fx(t))]
    # settling scales the transfer; maxdiff sets the threshold
    assert cascade(S, 0, 0, settling=0.5) == [(1, O.fx(F32(F32(0.5) * excess) / F32(2.0)))]
    assert cascade(S, 0, 0, maxdiff=0.2) == []
    # the lower cell receives through its own event as well (sign flips)
    got = cascade(S, 0, 1)
    assert got == [(0, -O.fx(t))]


def test_diagonal_distance_is_sqrt2():
    S = np.array([[1.0, 1.0], [1.0, 0.0]], np.float32)
    got = cascade(S, 0, 0)
    d = np.sqrt(F32(2.0))
    excess = F32(1.0) - F32(d * F32(0.005))
    t = excess / F32(2.0)
    assert got == [(3, O.fx(t))]


def test_tie_order_follows_offset_order_with_evolving_centre():
    # centre (1,1) = 1; zeros at offsets (-1,0) [cell (0,1)] and (0,1) [cell (1,2)]; others equal
    S = np.ones((3, 3), np.float32)
    S[0, 1] = 0.0
    S[1, 2] = 0.0
    got = cascade(S, 1, 1)
    e1 = F32(1.0) - F32(0.005)
    t1 = e1 / F32(2.0)
    hc = F32(1.0) - t1
    e2 = hc - F32(0.005)
    t2 = e2 / F32(2.0)
    assert got == [(1, O.fx(t1)), (5, O.fx(t2))]
    assert t1 != t2


def test_higher_neighbour_first_then_evolved_centre_reaches_equal_ones():
    # centre 0.5, neighbour (0,0) [offset (-1,-1), diagonal] at 1.0; the others at 0.5
    S = np.full((3, 3), 0.5, np.float32)
    S[0, 0] = 1.0
    got = cascade(S, 1, 1)
    d = np.sqrt(F32(2.0))
    e = F32(F32(0.5) - F32(1.0))
    t0 = (abs(e) - d * F32(0.005)) / F32(2.0)
    hc = F32(0.5) + t0
    assert got[0] == (0, -O.fx(t0))
    # after gaining t0 the centre exceeds its equal neighbours by t0 > d*maxdiff:
    # they follow in offset order, each with the centre still evolving
    order = [1, 2, 3, 5, 6, 7, 8]
    dist = {1: 1.0, 2: d, 3: 1.0, 5: 1.0, 6: d, 7: 1.0, 8: d}
    want = [(0, -O.fx(t0))]
    for c in order:
        diff = hc - F32(0.5)
        ex = abs(diff) - F32(F32(dist[c]) * F32(0.005))
        if ex <= 0:
            continue
        t = ex / F32(2.0)
        hc = hc - t
        want.append((c, O.fx(t)))
    assert got == want
    assert len(got) >= 2


def test_no_transfer_below_threshold_and_flat():
    S = np.full((3, 3), 0.25, np.float32)
    assert cascade(S, 1, 1) == []
    S[1, 1] = 0.25 + 0.004
    assert cascade(S, 1, 1) == []


def test_sediment_layer_caps_transfer():
    S = np.array([[0.1, 0.0]], np.float32)
    SS = np.array([[0.01, 0.0]], np.float32)
    got = cascade(S, 0, 0, SS=SS, layers=2)
    assert got == [(1, O.fx(F32(0.01)))]
    SS[0, 0] = 0.0
    assert cascade(S, 0, 0, SS=SS, layers=2) == []


def test_periodic_neighbours_wrap():
    S = np.zeros((4, 4), np.float32)
    S[0, 0] = 1.0
    got = dict(cascade(S, 0, 0, periodic=1))
    # eight neighbours, all lower: the first (offset (-1,-1)) is cell (3,3)
    assert len(got) == 8
    assert set(got) == {15, 12, 13, 3, 1, 7, 4, 5}


@pytest.mark.parametrize("settling,events", [(1.0, 3), (0.5, 7), (0.1, 40), (0.01, 100), (0.01, 1000)])
def test_repeated_events_on_isolated_pair_follow_geometric_law(settling, events):
    """V events at one cell of an isolated pair: consecutive events each remove
    `settling` of the remaining excess, so V of them move
    (a - maxdiff)(1 - (1-settling)^V)/2 -- the window's cascade uses exactly
    that settling factor."""
    a = 0.3
    S = np.array([[a, 0.0]], np.float32)
    got = cascade(S, 0, 0, events=events, settling=settling)
    moved = O.unfx(got[0][1])
    want = (a - 0.005) * (1 - (1 - settling) ** events) / 2
    assert abs(moved - want) < 2e-7 * (1 + np.log2(events))
    # one event equals the single-event formula exactly
    one = cascade(S, 0, 0, events=1, settling=settling)
    assert one[0][1] == O.fx(F32(F32(settling) * (F32(a) - F32(0.005))) / F32(2.0))


def test_settle_factor_by_binary_powering():
    lib = O.load()
    assert lib.or_settle(0.01, 1) == F32(0.01)
    assert lib.or_settle(0.01, 0) == F32(0.01)
    b = F32(1.0) - F32(0.01)
    # 5 = 101b: r = b * (b^2)^2, least significant bit first
    assert lib.or_settle(0.01, 5) == F32(1.0) - F32(b * F32(F32(b * b) * F32(b * b)))
    assert lib.or_settle(1.0, 7) == 1.0
    for v in (2, 3, 17, 255, 4096):
        assert abs(lib.or_settle(0.01, v) - (1 - 0.99 ** v)) < 1e-5
