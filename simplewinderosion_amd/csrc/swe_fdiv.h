/* swe_fdiv.h -- float32 division by the constant 500 without a divide
 * instruction sequence, for the turbulence term of Wind::fly()
 * (wind.hpp:83, glm's vec3 / float).
 *
 * q = RN(x * c) with c = RN(1/500), then one correction by the exact FMA
 * residual r = x - 500 q.  For every float x with 2^-100 <= |x| <= FLT_MAX
 * this returns the correctly rounded x / 500 (checked over all 2^32 inputs
 * by tests/test_fdiv.py); smaller |x| (the quotient is denormal or close to
 * it, where the residual is not exact), infinities and NaN take the
 * division.  Plain C so that the same text compiles for the host check and
 * for the device (where fmaf is v_fma_f32 with IEEE denormals). */
#ifndef SWE_FDIV_H
#define SWE_FDIV_H

#if defined(__HIPCC__)
#define SWE_FDIV_HD __host__ __device__
#else
#define SWE_FDIV_HD
#endif

static inline SWE_FDIV_HD float swe_div500(float x) {
  const float ax = __builtin_fabsf(x);
  if (__builtin_expect(ax >= 0x1p-100f && ax <= 3.40282346638528859812e+38f, 1)) {
    const float c = 0x1.0624dep-9f; /* RN(1/500) */
    const float q = x * c;
    const float r = __builtin_fmaf(-q, 500.0f, x);
    return __builtin_fmaf(r, c, q);
  }
  return x / 500.0f;
}

#endif
