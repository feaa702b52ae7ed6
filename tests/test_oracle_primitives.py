"""Pins of the oracle's scalar primitives against mathematics, not against itself.

fixed point <-> float: exact rational rounding (fractions.Fraction);
expf: error bound against double-precision exp;
RNG: the published SplitMix64 output sequence.
"""
import math
import random
from fractions import Fraction

import numpy as np
import pytest

import oracle as O


def rne(fr):
    """Round a Fraction to the nearest integer, ties to even."""
    fl = math.floor(fr)
    rem = fr - fl
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and fl % 2 == 1):
        return fl + 1
    return fl


def f32(x):
    return float(np.float32(x))


def test_fx_matches_exact_rounding():
    lib = O.load()
    rng = random.Random(1)
    vals = [0.0, -0.0, 1.0, -1.0, 0.5, 2.0 ** -32, 2.0 ** -33, 3 * 2.0 ** -33, 5 * 2.0 ** -34,
            -(2.0 ** -33), -3 * 2.0 ** -33, 1e-10, -1e-10, 123.456, -7.75, 2.0 ** 30, 1e-30]
    for _ in range(20000):
        e = rng.randint(-60, 30)
        vals.append(f32(rng.uniform(-1, 1) * 2.0 ** e))
    for v in vals:
        v = f32(v)
        want = rne(Fraction(v) * 2 ** 32)
        assert lib.or_fx(v) == want, v


def test_fx_ties_and_saturation():
    lib = O.load()
    assert lib.or_fx(f32(2.0 ** -33)) == 0            # 0.5 -> 0 (even)
    assert lib.or_fx(f32(3 * 2.0 ** -33)) == 2        # 1.5 -> 2
    assert lib.or_fx(f32(5 * 2.0 ** -33)) == 2        # 2.5 -> 2
    assert lib.or_fx(f32(-3 * 2.0 ** -33)) == -2
    assert lib.or_fx(float("inf")) == 2 ** 63 - 1
    assert lib.or_fx(float("-inf")) == -(2 ** 63)
    assert lib.or_fx(float("nan")) == 0
    assert lib.or_fx(f32(2.0 ** 40)) == 2 ** 63 - 1


def test_unfx_matches_exact_rounding():
    lib = O.load()
    rng = random.Random(2)
    vals = [0, 1, -1, 2 ** 32, -(2 ** 32), 2 ** 24 + 1, (2 ** 25 + 1) * 2 ** 10, (2 ** 25 + 3), 2 ** 62, -(2 ** 63)]
    for _ in range(20000):
        vals.append(rng.randint(-(2 ** rng.randint(1, 63)), 2 ** rng.randint(1, 63)))
    for m in vals:
        got = lib.or_unfx(m)
        # nearest float32 to m*2^-32, ties to even, via the exact rational
        fr = Fraction(m, 2 ** 32)
        want = f32(float(fr))  # float() of a Fraction is correctly rounded to double
        # double rounding guard: verify against exact comparison of neighbours
        w32 = np.float32(want)
        lo, hi = np.nextafter(w32, np.float32(-np.inf)), np.nextafter(w32, np.float32(np.inf))
        dw = abs(Fraction(float(w32)) - fr)
        assert dw <= abs(Fraction(float(lo)) - fr) and dw <= abs(Fraction(float(hi)) - fr)
        assert np.float32(got) == w32 or (abs(Fraction(float(got)) - fr) == dw and
                                          int(np.float32(got).view(np.uint32)) % 2 == 0), m


def test_vectorised_fx_helpers_agree_with_scalar():
    rng = np.random.default_rng(3)
    a = (rng.standard_normal(5000) * 10.0 ** rng.integers(-12, 6, 5000)).astype(np.float32)
    lib = O.load()
    assert np.array_equal(O._fx_vec(a), np.array([lib.or_fx(float(v)) for v in a], np.int64))
    m = rng.integers(-(2 ** 40), 2 ** 40, 5000)
    assert np.array_equal(O.unfx_vec(m), np.array([lib.or_unfx(int(v)) for v in m], np.float32))


def test_expf_error_bound():
    lib = O.load()
    xs = np.concatenate([np.linspace(-104, 0, 100003, dtype=np.float32),
                         np.linspace(-1, 1, 20001, dtype=np.float32),
                         np.linspace(0, 88, 10001, dtype=np.float32)])
    worst = 0.0
    for v in xs:
        got = lib.or_expf(float(v))
        t = math.exp(float(v))
        if t < 2.0 ** -126:
            # subnormal range: absolute error within one subnormal step
            assert abs(got - t) <= 2.0 ** -149
            continue
        ulp = float(np.spacing(np.float32(t)))
        worst = max(worst, abs(got - t) / ulp)
    assert worst <= 2.0, worst
    assert lib.or_expf(0.0) == 1.0
    assert lib.or_expf(-105.0) == 0.0
    assert math.isinf(lib.or_expf(89.0))


def test_splitmix64_published_sequence():
    # SplitMix64 (Steele, Lea, Flood 2014), state 1234567, increment then mix
    lib = O.load()
    x = 1234567
    got = []
    for _ in range(5):
        x = (x + 0x9E3779B97F4A7C15) % 2 ** 64
        got.append(lib.or_mix64(x))
    assert got == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                   4593380528125082431, 16408922859458223821]


def test_spawn_and_turbulence_ranges():
    import ctypes
    lib = O.load()
    px, pz = ctypes.c_float(), ctypes.c_float()
    r = (ctypes.c_int32 * 3)()
    cells = set()
    for i in range(4000):
        k = lib.or_particle_key(42, 3, i)
        lib.or_spawn(k, 7, 5, ctypes.byref(px), ctypes.byref(pz))
        assert 0.0 <= px.value < 7.0 and 0.0 <= pz.value < 5.0
        # fractional part is a multiple of 1/1000 as world.hpp:54 builds it
        fx = round((px.value - math.floor(px.value)) * 1000)
        assert abs(px.value - (math.floor(px.value) + f32(fx / 1000.0))) < 1e-6
        cells.add((int(px.value), int(pz.value)))
        lib.or_turb(k, i % 513, r)
        assert all(0 <= r[j] <= 1000 for j in range(3))
    assert len(cells) == 35
