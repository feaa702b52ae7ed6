// swe_device.hpp -- scalar device primitives of the erosion hot path.
//
// Every function here is a sequence of IEEE-754 float32 operations (plus one
// float64 multiply/subtract in the drag factor, wind.hpp:87) written in the
// operation order of the reference and of glm 0.9.9, so that with
// -ffp-contract=off and correctly-rounded division/sqrt the device produces
// the same bits as any other IEEE implementation of the same steps.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/swe.h"

namespace swe {

// ---------------------------------------------------------------------------
// Bits and fixed point (int64, resolution 2^-32)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float u2f(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ uint32_t f2u(float f) { return __float_as_uint(f); }

// round_half_even(v * 2^32), computed on the IEEE encoding. NaN -> 0,
// overflow saturates.  The general (slow) form; fx_from_f32 below takes it
// only for |v| >= 2^31 and NaN.
__device__ __forceinline__ int64_t fx_from_f32_bits(float v) {
  const uint32_t u = f2u(v);
  const uint32_t e = (u >> 23) & 0xffu;
  const uint64_t sgn = u >> 31;
  if (e == 0xffu) {
    if (u & 0x7fffffu) return 0;
    return sgn ? INT64_MIN : INT64_MAX;
  }
  if (e == 0) return 0;
  const uint64_t m = (u & 0x7fffffu) | 0x800000u;
  const int sh = (int)e - 118;
  uint64_t r;
  if (sh >= 0) {
    if (sh > 39) return sgn ? INT64_MIN : INT64_MAX;
    r = m << sh;
    if (r > (uint64_t)INT64_MAX) return sgn ? INT64_MIN : INT64_MAX;
  } else {
    const int k = -sh;
    if (k > 25) return 0;
    const uint64_t q = m >> k;
    const uint64_t rem = m & ((1ull << k) - 1ull);
    const uint64_t half = 1ull << (k - 1);
    r = q + ((rem > half || (rem == half && (q & 1ull))) ? 1ull : 0ull);
  }
  return sgn ? -(int64_t)r : (int64_t)r;
}

// Same result for every input: for |v| < 2^31, v * 2^32 is exact and below
// 2^63, rint rounds it half-to-even to an integral float, and the integral
// float converts to int64 exactly (denormal v give 0 either way).
__device__ __forceinline__ int64_t fx_from_f32(float v) {
  if (__builtin_expect(__builtin_fabsf(v) < 2147483648.0f, 1))
    return (int64_t)__builtin_rintf(v * 4294967296.0f);
  return fx_from_f32_bits(v);
}

// round_half_even(m * 2^-32) to float32: the int64 -> float32 conversion
// rounds to nearest even, and the power-of-two scale is exact (|m| >= 1
// keeps the result far from the denormal range).
__device__ __forceinline__ float f32_from_fx(int64_t m) { return (float)m * 2.3283064365386963e-10f; }

// The same by integer operations (reference form of f32_from_fx).
__device__ __forceinline__ float f32_from_fx_bits(int64_t m) {
  if (m == 0) return 0.0f;
  const bool neg = m < 0;
  const uint64_t a = neg ? (uint64_t)0 - (uint64_t)m : (uint64_t)m;
  const int msb = 63 - __clzll((long long)a);
  uint64_t q;
  int ex;
  if (msb <= 23) {
    q = a;
    ex = -32;
  } else {
    const int k = msb - 23;
    q = a >> k;
    const uint64_t rem = a & ((1ull << k) - 1ull);
    const uint64_t half = 1ull << (k - 1);
    if (rem > half || (rem == half && (q & 1ull))) q += 1;
    ex = k - 32;
  }
  const float r = (float)(uint32_t)q * u2f((uint32_t)(ex + 127) << 23);
  return neg ? -r : r;
}

// exp for float32: Cody-Waite reduction with a two-part ln2, degree-7
// Taylor polynomial in Horner form, exact power-of-two scaling (two steps
// below 2^-125).
__device__ __forceinline__ float exp_f32(float x) {
  if (x != x) return x;
  if (x > 88.72283935546875f) return u2f(0x7f800000u);
  if (x < -104.0f) return 0.0f;
  const float log2e = 1.44269502162933349609375f;
  const float ln2_hi = 0.693145751953125f;
  const float ln2_lo = 1.428606765330187045e-06f;
  const float kf = __builtin_rintf(x * log2e);
  const float r = (x - kf * ln2_hi) - kf * ln2_lo;
  float p = 1.0f / 5040.0f;
  p = p * r + 1.0f / 720.0f;
  p = p * r + 1.0f / 120.0f;
  p = p * r + 1.0f / 24.0f;
  p = p * r + 1.0f / 6.0f;
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  int k = (int)kf;
  if (k < -125) {
    p = p * u2f((uint32_t)(127 - 64) << 23);
    k += 64;
  }
  return p * u2f((uint32_t)(k + 127) << 23);
}

// ---------------------------------------------------------------------------
// Counter RNG (SplitMix64 output function over (seed, cycle, particle, call))
// ---------------------------------------------------------------------------
constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;
constexpr uint64_t kPart = 0xD1B54A32D192ED03ull;

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ __forceinline__ uint64_t cycle_key(uint64_t seed, uint64_t cycle) {
  return mix64(seed + kGolden * (cycle + 1));
}
__host__ __device__ __forceinline__ uint64_t particle_key(uint64_t ck, uint64_t particle) {
  return mix64(ck ^ (kPart * (particle + 1)));
}

// ---------------------------------------------------------------------------
// glm vec3 arithmetic (operation order of glm 0.9.9)
// ---------------------------------------------------------------------------
struct f3 {
  float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 mul(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ f3 smul(float s, f3 a) { return mk(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ f3 muls(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ f3 divs(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ f3 subs(f3 a, float s) { return mk(a.x - s, a.y - s, a.z - s); }
__device__ __forceinline__ float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
  return mk(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
__device__ __forceinline__ float length(f3 a) { return __builtin_sqrtf(dot(a, a)); }
__device__ __forceinline__ f3 normalize(f3 a) { return muls(a, 1.0f / __builtin_sqrtf(dot(a, a))); }
__device__ __forceinline__ f3 mix(f3 x, f3 y, float a) { return add(muls(x, 1.0f - a), muls(y, a)); }

// quad::_normal (cellpool.h:200-223) from the five stencil heights.  gNE,
// gSW, gSE, gNW gate the four planes (diagonal in bounds).
__device__ __forceinline__ f3 surface_normal(float hp, float hyp, float hxp, float hym, float hxm,
                                             bool g1, bool g2, bool g3, bool g4, float mapscale) {
  const f3 s = mk(1.0f, mapscale, 1.0f);
  f3 n = mk(0.0f, 0.0f, 0.0f);
  if (g1) n = add(n, cross(mul(s, mk(0.0f, hyp - hp, 1.0f)), mul(s, mk(1.0f, hxp - hp, 0.0f))));
  if (g2) n = add(n, cross(mul(s, mk(0.0f, hym - hp, -1.0f)), mul(s, mk(-1.0f, hxm - hp, 0.0f))));
  if (g3) n = add(n, cross(mul(s, mk(1.0f, hxp - hp, 0.0f)), mul(s, mk(0.0f, hym - hp, -1.0f))));
  if (g4) n = add(n, cross(mul(s, mk(-1.0f, hxm - hp, 0.0f)), mul(s, mk(0.0f, hyp - hp, 1.0f))));
  if (length(n) > 0.0f) n = normalize(n);
  return n;
}

// Euclidean remainder of v by n (n > 0), exact for all int32 v: a float
// reciprocal estimate corrected by at most one step, integer fallback for
// |v| >= 2^23.
__device__ __forceinline__ int32_t wrap(int32_t v, int32_t n, float rcp_n) {
  if ((uint32_t)v < (uint32_t)n) return v;
  if (v > -(1 << 23) && v < (1 << 23)) {
    int32_t q = (int32_t)__builtin_floorf((float)v * rcp_n);
    int32_t r = v - q * n;
    if (r < 0) r += n;
    if (r >= n) r -= n;
    return r;
  }
  int32_t r = v % n;
  return r < 0 ? r + n : r;
}

}  // namespace swe
