"""Pins of one fly() call (wind.hpp:40-112) off the flat, grounded map.

Each test drives the oracle's one-step export (or_fly_once: a particle in an
arbitrary state against a frozen map) and checks a line of wind.hpp against
something the oracle does not compute itself: an exact special case, a
vector identity, libm's exp, or a float64 model of the step written from the
reference's formulas with identities in place of glm's cross products.  The
batched and the sequential restatements share the arithmetic of :56-106, so
these pins cover both; the sequential-only parts (glibc draws, in-place
pickup, World::cascade on a float map, the normal) have their own pins below.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle as O

F32 = np.float32


def _turb(key, age):
    out = (ctypes.c_int32 * 3)()
    O.load().or_turb(ctypes.c_uint64(key), int(age), out)
    return np.array(list(out), np.float64)


def _ulp(x):
    x = abs(F32(x))
    return float(np.nextafter(x, F32(np.inf)) - x)


def _state(px, py, pz, v, sed):
    return np.array([px, py, pz, v[0], v[1], v[2], sed], np.float32)


def _flat(h=0.0, n=32):
    return np.full((n, n), h, np.float32)


def _plane(a, b, c0=0.5, n=64):
    """h = a*x + b*y + c0 with power-of-two slopes: every height is exact, so
    quad::_normal is normalize(-mapscale*a, 1, -mapscale*b) on interior cells."""
    x = np.arange(n, dtype=np.float64)[:, None]
    y = np.arange(n, dtype=np.float64)[None, :]
    return (a * x + b * y + c0).astype(np.float32)


# ---------------------------------------------------------------------------
# wind.hpp:87 -- the drag factor is (float)(1.0 - 0.3*(double)sediment)
# ---------------------------------------------------------------------------
def _drag_cases():
    """Sediment values where the double-precision factor and the all-float
    formula 1.0f - 0.3f*s round to different floats."""
    rng = np.random.default_rng(5)
    s = np.concatenate([rng.random(4000).astype(np.float32) * F32(3.0), np.float32([0.05, 0.3, 1.0, 2.5])])
    dbl = (1.0 - 0.3 * s.astype(np.float64)).astype(np.float32)
    flt = (F32(1.0) - F32(0.3) * s).astype(np.float32)
    differ = s[dbl != flt]
    same = s[dbl == flt]
    return np.concatenate([differ[:24], same[:4]]), len(differ)


def test_drag_factor_is_double_precision():
    """Grounded on a flat map with speed == pspeed == (1,0,0), every term of
    :61-83 leaves the speed at exactly (1,0,0) (shadow 1, hfac 1, collision
    -0, rspeed == speed, mix(.., 1) == rspeed, turbulence scaled by -0), so
    the new speed.x is the drag factor itself."""
    cases, ndiffer = _drag_cases()
    assert ndiffer > 100  # the float formula really differs on this set
    p = O.reference_params(periodic=1)
    S = _flat(0.25)
    for s in cases:
        ok, st, age, tr, _ = O.fly_once(p, S, _state(8.2, 0.25, 9.7, (1, 0, 0), s), 3, key=11)
        assert ok and age == 4
        want = F32(1.0 - 0.3 * float(s))
        assert tr["drag"] == want, s
        assert st[3] == want and st[4] == 0 and st[5] == 0, s
        assert np.array_equal(tr["sp_terrain"], np.float32([1, 0, 0]))


# ---------------------------------------------------------------------------
# wind.hpp:50-52, 65-66, 83 -- an airborne particle over a flat map
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("sed,dy,key", [(1.5, 0.5, 3), (0.8, 0.1, 17), (2.2, 1.7, 99), (0.3, 0.02, 1234)])
def test_airborne_step_closed_form(sed, dy, key):
    """n = (0,1,0), speed == pspeed == (1,0,0), pos.y = h + dy, age >= 1.
    Then shadow = 1 and :61 leaves the speed unchanged; gravity (:66) makes
    it (1, -g*s, 0); collision c = g*s/sqrt(1+(g*s)^2); rspeed = (1-c)(1,0,0);
    :79 gives (1 - 0.9*c*hfac, -0.1*g*s, 0); :83 adds
    0.1*hfac*c*(r-500)/500 with the call's three draws; :87 scales by
    1 - 0.3*s."""
    g, bl = 0.1, 2.0
    p = O.reference_params(periodic=1)
    h = 0.25
    S = _flat(h)
    py = F32(h) + F32(dy)
    age = 7
    ok, st, age2, tr, pick = O.fly_once(p, S, _state(10.4, py, 11.6, (1, 0, 0), sed), age, key)
    assert ok and age2 == age + 1
    gs = F32(F32(g) * F32(sed))
    # :65-66 -- exactly one float subtraction from speed.y before the later terms
    assert np.array_equal(tr["sp_gravity"], np.float32([1.0, F32(0.0) - gs, 0.0]))
    # :50 -- hfac from the clamped height difference
    x = F32(-(F32(py - F32(h)))) / F32(bl)
    hfac = math.exp(float(x))
    assert abs(float(tr["hfac"]) - hfac) <= 2 * _ulp(hfac)
    c = float(gs) / math.sqrt(1.0 + float(gs) ** 2)
    assert abs(float(tr["collision"]) - c) <= 1e-6
    r = _turb(key, age)
    assert np.array_equal(tr["draws"], r.astype(np.float32))
    turb = 0.1 * hfac * c * (r - 500.0) / 500.0
    v = np.array([1.0 - 0.9 * c * hfac, -0.1 * float(gs), 0.0]) + turb
    v *= 1.0 - 0.3 * float(sed)
    assert np.allclose(st[3:6], v, rtol=2e-6, atol=2e-7), (st[3:6], v)
    # the move (:91) and no snap to the ground while airborne (:42-43)
    assert np.allclose(st[:3], np.array([10.4, float(py), 11.6]) + v, rtol=0, atol=4e-6)
    # the pickup (:106-112): force = max(0, -v.n), lift = (1-c)|v|
    vn = float(np.linalg.norm(v))
    cap = max(0.0, -v[1]) * hfac + 0.02 * (1 - c) * vn * hfac
    q = 0.05 * (cap - sed)
    assert abs(float(st[6]) - (sed + q)) <= 2e-6 * max(1.0, sed)
    assert abs(pick / 2.0 ** 32 - q) <= 1e-6 * abs(q) + 1e-9


@pytest.mark.parametrize("dy", [0.0, 1e-3, 0.07, 0.5, 1.0, 3.0, 9.5, 40.0])
def test_hfac_is_expf_of_the_height_above_ground(dy):
    """:50-52 -- hfac = exp(-(pos.y - h)/boundarylayer), within 2 ulp of the
    exact exponential; 1 exactly on the ground."""
    p = O.reference_params(periodic=1, boundarylayer=2.0)
    h = 0.375
    py = F32(h) + F32(dy)
    _, _, _, tr, _ = O.fly_once(p, _flat(h), _state(4.0, py, 5.0, (0.3, 0, 0.2), 0.1), 2, key=8)
    x = F32(-(F32(py - F32(h)))) / F32(2.0)
    want = math.exp(float(x))
    assert abs(float(tr["hfac"]) - want) <= 2 * _ulp(want)
    if dy == 0.0:
        assert tr["hfac"] == 1.0


def test_fresh_particle_snaps_to_the_ground_and_below_ground_is_lifted():
    """:42-43 -- age 0 or pos.y below the cell: pos.y = h, so hfac = 1 and no gravity."""
    p = O.reference_params(periodic=1)
    S = _flat(0.6)
    for age, py in ((0, 5.0), (4, 0.1)):
        _, _, _, tr, _ = O.fly_once(p, S, _state(6.0, py, 6.0, (0.2, 0.0, 0.1), 1.0), age, key=2)
        assert tr["hfac"] == 1.0
        assert tr["sp_gravity"][1] == tr["sp_prevailing"][1]


# ---------------------------------------------------------------------------
# wind.hpp:56-59, 75 -- shadow and the redirect on tilted planes
# ---------------------------------------------------------------------------
SLOPES = [(1 / 256, 0.0), (-1 / 256, 0.0), (0.0, 1 / 128), (1 / 64, -1 / 512), (-1 / 32, 1 / 32), (1 / 1024, 1 / 1024)]


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("a,b", SLOPES)
def test_redirect_is_tangential_and_shadow_follows_the_normal(a, b):
    """rspeed = n x ((1-c) v x n) is (1-c)(v - (v.n) n) for a unit normal
    (the vector triple product), so rspeed.n == 0; a transposed cross product
    flips its sign.  shadow = 1 - max(0, normalize(pspeed).n)."""
    S = _plane(a, b)
    n_exact = _unit([-80.0 * a, 1.0, -80.0 * b])
    rng = np.random.default_rng(int(a * 4096) + 7 * int(b * 4096) + 100)
    ps = _unit([1.0, 0.0, 1.0]).astype(np.float32)
    p = O.reference_params(periodic=0, pspeed=ps)
    for i in range(40):
        x, y = rng.uniform(8, 55, 2)
        hc = float(S[int(np.floor(x + 0.5)), int(np.floor(y + 0.5))])
        v = rng.uniform(-0.6, 0.6, 3)
        airborne = i % 2
        py = F32(hc + (rng.uniform(0.01, 2.0) if airborne else 0.0))
        _, _, _, tr, _ = O.fly_once(p, S, _state(x, py, y, v, rng.uniform(0, 1.5)), 5, key=i + 1)
        n = tr["n"].astype(np.float64)
        assert np.allclose(n, n_exact, atol=2e-7)
        shadow = 1.0 - max(0.0, float(np.dot(_unit(ps), n_exact)))
        assert abs(float(tr["shadow"]) - shadow) <= 2e-7
        vg = tr["sp_gravity"].astype(np.float64)
        c = float(tr["collision"])
        rs = tr["rspeed"].astype(np.float64)
        want = (1.0 - c) * (vg - np.dot(vg, n) * n)
        scale = np.linalg.norm(vg) + 1e-30
        assert abs(np.dot(rs, n)) <= 4e-7 * scale
        assert np.allclose(rs, want, atol=4e-7 * scale)
        # :70-71: collision = max(0, -normalize(v).n)
        assert abs(c - max(0.0, -np.dot(vg, n) / np.linalg.norm(vg))) <= 3e-7


# ---------------------------------------------------------------------------
# The whole call against a float64 model written from wind.hpp:42-112
# ---------------------------------------------------------------------------
def _model(p, h, n, st, age, draws):
    """One fly() call in float64: wind.hpp:42-112 with n x (a x n) = a - (a.n)n
    and -normalize(v).n * |v| = -v.n."""
    pos = st[:3].astype(np.float64)
    v = st[3:6].astype(np.float64)
    s = float(st[6])
    h = float(h)
    if age == 0 or pos[1] < h:
        pos[1] = h
    hfac = max(0.0, math.exp(-(pos[1] - h) / p.boundarylayer))
    ps = np.array(list(p.pspeed), np.float64)
    shadow = 1.0 - max(0.0, float(np.dot(_unit(ps), n)))
    v = v + 0.05 * ((0.1 + 0.9 * shadow) * ps - v)
    if pos[1] > h:
        v[1] -= p.gravity * s
    c = max(0.0, -float(np.dot(v, n)) / float(np.linalg.norm(v)))
    rs = (1.0 - c) * (v - float(np.dot(v, n)) * n)
    v = v + 0.9 * (shadow * ((1.0 - shadow * hfac) * ps + shadow * hfac * rs) - v)
    v = v + 0.1 * hfac * c * (draws - 500.0) / 500.0
    v = v * (1.0 - 0.3 * s)
    pos = pos + v
    force = max(0.0, -float(np.dot(v, n)))
    lift = (1.0 - c) * float(np.linalg.norm(v))
    cap = force * hfac + 0.02 * lift * hfac
    q = p.suspension * (cap - s)
    return pos, v, s + q, q


@pytest.mark.parametrize("a,b", SLOPES + [(0.0, 0.0)])
@pytest.mark.parametrize("wind", ["x", "diag", "elevated"])
def test_one_step_matches_float64_model(a, b, wind):
    ps = {"x": np.float32([1, 0, 0]), "diag": _unit([1, 0, 1]).astype(np.float32),
          "elevated": _unit([0.6, 0.4, -0.5]).astype(np.float32)}[wind]
    p = O.reference_params(periodic=0, pspeed=ps, suspension=0.05)
    S = _plane(a, b)
    n_exact = _unit([-80.0 * a, 1.0, -80.0 * b])
    rng = np.random.default_rng(hash((a, b, wind)) % (2 ** 32))
    seen_turb = 0
    for i in range(60):
        x, y = rng.uniform(8, 55, 2)
        hc = F32(S[int(np.floor(x + 0.5)), int(np.floor(y + 0.5))])
        airborne = i % 3 != 0
        py = F32(hc + F32(rng.uniform(0.01, 2.5))) if airborne else hc
        v = rng.uniform(-0.8, 0.8, 3)
        sed = rng.uniform(0.0, 2.0)
        age = int(rng.integers(1, 500))
        key = int(rng.integers(1, 2 ** 62))
        st0 = _state(x, py, y, v, sed)
        ok, st, _, tr, pick = O.fly_once(p, S, st0, age, key)
        assert ok
        draws = _turb(key, age)
        pos, vm, sm, q = _model(p, hc, n_exact, st0, age, draws)
        vs = max(1.0, float(np.linalg.norm(vm)))
        assert np.allclose(st[3:6], vm, rtol=0, atol=2e-5 * vs), (i, st[3:6], vm)
        assert np.allclose(st[:3], pos, rtol=0, atol=2e-5 * vs + 8e-6), (i, st[:3], pos)
        assert abs(float(st[6]) - sm) <= 2e-5 * max(1.0, sm), i
        assert abs(pick / 2.0 ** 32 - q) <= 2e-5 * max(abs(q), 1e-3), i
        if float(tr["collision"]) * float(tr["hfac"]) > 0.01:
            seen_turb += 1
    assert seen_turb >= 5  # turbulence really acts on this set


def test_bounded_particle_leaving_the_map_dies_and_periodic_wraps():
    """:32-38 -- a rounded position outside a bounded map ends the particle;
    the periodic variant reads the wrapped cell (cellpool.h:96-113,447-465)."""
    S = _plane(1 / 256, 0.0, n=32)
    p = O.reference_params(periodic=0)
    ok, _, _, _, _ = O.fly_once(p, S, _state(31.6, 0.5, 4.0, (1, 0, 0), 0.0), 3, key=4)
    assert not ok
    pp = O.reference_params(periodic=1)
    ok1, st1, _, tr1, _ = O.fly_once(pp, S, _state(31.6 + 32.0, 5.0, 4.0 - 64.0, (1, 0, 0), 0.2), 3, key=4)
    ok2, st2, _, tr2, _ = O.fly_once(pp, S, _state(31.6 - 32.0, 5.0, 4.0, (1, 0, 0), 0.2), 3, key=4)
    assert ok1 and ok2
    assert tr1["h"] == S[0, 4] and tr2["h"] == S[0, 4]
    assert np.array_equal(st1[3:7], st2[3:7])


# ---------------------------------------------------------------------------
# The sequential restatement (the reference's own order of events)
# ---------------------------------------------------------------------------
def _libc():
    return ctypes.CDLL("libc.so.6")


def test_sequential_draws_are_glibc_rand_mod_1001():
    """wind.hpp:83 -- three rand()%1001 draws per call, in order."""
    libc = _libc()
    p = O.reference_params(periodic=1)
    S = _plane(1 / 64, 1 / 128)
    for seed in (1, 42, 2024):
        _, _, _, tr, _, _ = O.fly_seq_once(p, S, _state(9.0, 0.0, 9.0, (0.1, -0.3, 0.0), 0.4), 3, seed)
        libc.srand(seed)
        want = [libc.rand() % 1001 for _ in range(3)]
        assert tr["draws"].tolist() == want


def test_sequential_step_shares_the_batched_arithmetic():
    """Same state, map and (grounded: hfac = 1 either way) the same stages up
    to the draws; the sequential normal (AoS map, bounds-checked lookups) is
    the batched one's bits, edge cells included."""
    rng = np.random.default_rng(3)
    S = (rng.random((24, 20)) * 0.3).astype(np.float32)
    for periodic in (0, 1):
        p = O.reference_params(periodic=periodic)
        for i in range(40):
            x = rng.uniform(-0.4, 23.4)
            y = rng.uniform(-0.4, 19.4)
            st0 = _state(x, 0.0, y, rng.uniform(-0.5, 0.5, 3), rng.uniform(0, 1))
            ok_b, _, _, tb, _ = O.fly_once(p, S, st0, 0, key=5)
            ok_s, _, _, ts, _, _ = O.fly_seq_once(p, S, st0, 0, seed=9)
            assert ok_b and ok_s
            for k in ("n", "h", "hfac", "shadow", "sp_prevailing", "sp_gravity", "collision", "rspeed",
                      "sp_terrain"):
                assert np.array_equal(tb[k], ts[k]), (periodic, i, k)


def test_sequential_pickup_is_written_into_the_cell():
    """:110-112 on the live map: height -= suspension*diff, sediment +=
    suspension*diff; with no slope above maxdiff the cascade moves nothing."""
    p = O.reference_params(periodic=1, suspension=0.05, maxdiff=10.0)
    S = _flat(0.3, n=16)
    for sed, v in ((0.0, (0.4, 0.0, 0.3)), (0.5, (1.0, 0.0, 0.0)), (0.01, (0.05, 0.0, -0.02))):
        ok, st, _, tr, H, _ = O.fly_seq_once(p, S, _state(7.3, 0.3, 8.6, v, sed), 2, seed=1)
        assert ok
        q = F32(F32(0.05) * tr["diff"])
        want = S.copy()
        want[7, 9] = F32(F32(0.3) - q)
        assert np.array_equal(H, want)
        assert st[6] == F32(F32(sed) + q)
        # on a flat map capacity = 0.02*|v|*hfac: diff has the sign of 0.02|v| - sed
        assert np.sign(tr["diff"]) == np.sign(0.02 * np.linalg.norm(st[3:6]) - sed)


def test_sequential_cascade_hand_cases():
    """World::cascade (world.hpp:79-153) on a float map in place: the pair
    transfer settling*(|dh| - d*maxdiff)/2, the diagonal distance sqrtf(2),
    equal neighbours in offset order with the centre evolving."""
    p = O.reference_params(periodic=0, settling=1.0, maxdiff=0.005)
    H, _ = O.cascade_seq(p, np.array([[1.0, 0.0]], np.float32), 0, 0)
    t = F32(F32(1.0) - F32(0.005)) / F32(2.0)
    assert np.array_equal(H, np.array([[F32(1.0) - t, t]], np.float32))
    H, _ = O.cascade_seq(p, np.array([[1.0, 1.0], [1.0, 0.0]], np.float32), 0, 0)
    td = F32(F32(1.0) - F32(np.sqrt(F32(2.0)) * F32(0.005))) / F32(2.0)
    assert H[1, 1] == td and H[0, 0] == F32(1.0) - td
    S = np.ones((3, 3), np.float32)
    S[0, 1] = 0.0
    S[1, 2] = 0.0
    H, _ = O.cascade_seq(p, S, 1, 1)
    t1 = F32(F32(1.0) - F32(0.005)) / F32(2.0)
    hc = F32(1.0) - t1
    t2 = F32(hc - F32(0.005)) / F32(2.0)
    assert H[0, 1] == t1 and H[1, 2] == t2 and H[1, 1] == F32(hc - t2)
    # the lower centre receives from a higher neighbour (diff < 0 branch)
    H, _ = O.cascade_seq(p, np.array([[0.0, 0.8]], np.float32), 0, 0)
    t = F32(F32(0.8) - F32(0.005)) / F32(2.0)
    assert np.array_equal(H, np.array([[t, F32(0.8) - t]], np.float32))
    # settling scales, maxdiff thresholds
    p2 = O.reference_params(periodic=0, settling=0.25, maxdiff=0.005)
    H, _ = O.cascade_seq(p2, np.array([[1.0, 0.0]], np.float32), 0, 0)
    t = F32(F32(0.25) * F32(F32(1.0) - F32(0.005))) / F32(2.0)
    assert H[0, 1] == t
    p3 = O.reference_params(periodic=0, maxdiff=0.6)
    H, _ = O.cascade_seq(p3, np.array([[0.5, 0.0]], np.float32), 0, 0)
    assert np.array_equal(H, np.array([[0.5, 0.0]], np.float32))
