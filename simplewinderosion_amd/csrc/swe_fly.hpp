// swe_fly.hpp -- one Wind::fly() call (wind.hpp:28-117) on the device.
//
// Shared by the global-memory march (k_march) and the tile-owned step
// kernel (k_step).  The map is read through `Map` (heights / sediment of
// the window's snapshot around a cell) and the particle's writes go through
// `Acc` (pickup deltas, cascade-event counts, flow tracks).
#pragma once

#include "swe_device.hpp"
#include "swe_fdiv.h"

namespace swe {

struct Particle {
  float px, py, pz, sed;
  float sx, sy, sz;
  int32_t age;     // < 0: dead
  int64_t ledger;  // fixed-point load carried
  int32_t lastc;   // pre-move cell of the last step (global index)
};

struct FlyConsts {
  int32_t nx, ny;
  float rcp_nx, rcp_ny;
  float mapscale;
  int32_t max_age;
  float boundarylayer, suspension, gravity, abrasion;
  f3 ps, nps;  // pspeed and normalize(pspeed)
  // boundarylayer a power of two with a normal reciprocal: dividing by it
  // and multiplying by rcp_bl round the same real number, so give one result
  bool bl_pow2;
  float rcp_bl;
};

// Fills bl_pow2 / rcp_bl from boundarylayer.
__device__ __forceinline__ void set_boundarylayer(FlyConsts& k, float bl) {
  k.boundarylayer = bl;
  const uint32_t u = f2u(bl), e = (u >> 23) & 0xffu;
  k.bl_pow2 = (u & 0x807fffffu) == 0u && e >= 1u && e <= 253u;
  k.rcp_bl = k.bl_pow2 ? u2f((254u - e) << 23) : 0.0f;
}

// Outcome of one call.
enum FlyOutcome { kStepped = 0, kDied = 1 };

// world.hpp:52-54,62 / wind.hpp:6-13 on the counter RNG; mode 1 is the
// commented boundary spawn of world.hpp:56-61 (x = 0 or y = 0 edge)
__device__ __forceinline__ void spawn(uint64_t pk, int32_t nx, int32_t ny, int32_t mode, Particle& w) {
  const uint64_t ra = mix64(pk + 1 * kGolden);
  if (mode == 1) {
    const uint32_t shift = (uint32_t)ra % (uint32_t)(nx + ny);
    w.px = shift < (uint32_t)nx ? (float)shift : 0.0f;
    w.pz = shift < (uint32_t)nx ? 0.0f : (float)(shift - (uint32_t)nx);
  } else {
    const uint64_t rb = mix64(pk + 2 * kGolden);
    const int32_t ix = (int32_t)((uint32_t)ra % (uint32_t)nx);
    const int32_t iy = (int32_t)((uint32_t)(ra >> 32) % (uint32_t)ny);
    const int32_t fxp = (int32_t)((uint32_t)rb % 1000u);
    const int32_t fyp = (int32_t)((uint32_t)(rb >> 32) % 1000u);
    w.px = (float)ix + (float)fxp / 1000.0f;
    w.pz = (float)iy + (float)fyp / 1000.0f;
  }
  w.py = 0.0f;
  w.sed = 0.0f;
  w.sx = w.sy = w.sz = 0.0f;
  w.age = 0;
  w.ledger = 0;
  w.lastc = 0;
}

// Cell of a position: ipos = round(vec2(pos.x, pos.z)) (wind.hpp:30), wrapped
// on a periodic map.  Returns false when the particle leaves a bounded map
// (wind.hpp:32-38) or its position is no longer a finite int32 pair.
template <bool PERIODIC>
__device__ __forceinline__ bool cell_of(float px, float pz, const FlyConsts& k, int32_t& cx, int32_t& cy) {
  const float rx = __builtin_roundf(px), rz = __builtin_roundf(pz);
  if (!(__builtin_fabsf(rx) < 1073741824.0f) || !(__builtin_fabsf(rz) < 1073741824.0f)) return false;
  const int32_t ix = (int32_t)rx, iy = (int32_t)rz;
  if (PERIODIC) {
    cx = wrap(ix, k.nx, k.rcp_nx);
    cy = wrap(iy, k.ny, k.rcp_ny);
    return true;
  }
  if (ix < 0 || iy < 0 || ix >= k.nx || iy >= k.ny) return false;
  cx = ix;
  cy = iy;
  return true;
}

// One fly() call for a particle whose current cell (cx, cy) is valid.
// `Map` provides h(dx, dy) -- snapshot height at (cx+dx, cy+dy) for the
// five-point stencil -- and s0() -- sediment snapshot at the cell.
// `Acc` receives add_pickup(cell, q_fx, both_layers), add_event(cell),
// add_track(cell, sed, speed), deposit(cell, ledger).
// Returns kDied when age > max_age (the caller handles the death).
template <bool PERIODIC, bool TRACK, int LAYERS, class Map, class Acc>
__device__ __forceinline__ FlyOutcome fly(Particle& w, uint64_t pk, int32_t cx, int32_t cy, const FlyConsts& k,
                                          const Map& map, Acc& acc) {
  const int64_t cell = (int64_t)cx * k.ny + cy;
  // five-point stencil of quad::_normal (cellpool.h:200-223)
  const float h = map.h(0, 0);
  float hyp = 0.0f, hym = 0.0f, hxp = 0.0f, hxm = 0.0f;
  bool g1, g2, g3, g4;
  if (PERIODIC) {
    hyp = map.h(0, 1);
    hym = map.h(0, -1);
    hxp = map.h(1, 0);
    hxm = map.h(-1, 0);
    g1 = g2 = g3 = g4 = true;
  } else {
    const bool bxp = cx + 1 < k.nx, bxm = cx > 0, byp = cy + 1 < k.ny, bym = cy > 0;
    g1 = bxp && byp;
    g2 = bxm && bym;
    g3 = bxp && bym;
    g4 = bxm && byp;
    if (byp) hyp = map.h(0, 1);
    if (bym) hym = map.h(0, -1);
    if (bxp) hxp = map.h(1, 0);
    if (bxm) hxm = map.h(-1, 0);
  }
  const f3 n = surface_normal(h, hyp, hxp, hym, hxm, g1, g2, g3, g4, k.mapscale);
  if (w.age == 0 || w.py < h) w.py = h;  // :42-43
  if (w.age > k.max_age) return kDied;  // :45-46
  const int32_t call = w.age;
  w.age = w.age + 1;

  const float hrel = -(w.py - h);
  float hfac = exp_f32(k.bl_pow2 ? hrel * k.rcp_bl : hrel / k.boundarylayer);  // :50-52
  if (hfac < 0.0f) hfac = 0.0f;

  float shadow = dot(k.nps, n);  // :56-59
  if (shadow < 0.0f) shadow = 0.0f;
  shadow = 1.0f - shadow;

  f3 sp = mk(w.sx, w.sy, w.sz);
  sp = add(sp, smul(0.05f, sub(muls(k.ps, 0.1f + 0.9f * shadow), sp)));  // :61

  if (w.py > h) sp.y -= k.gravity * w.sed;  // :65-66

  float collision = -dot(normalize(sp), n);  // :70-71
  if (collision < 0.0f) collision = 0.0f;

  const f3 rspeed = cross(n, cross(smul(1.0f - collision, sp), n));  // :75

  sp = add(sp, smul(0.9f, sub(smul(shadow, mix(k.ps, rspeed, shadow * hfac)), sp)));  // :79

  {  // :83 turbulence, three draws in [0,1000] for this call
    const uint64_t t = mix64(pk + (uint64_t)(call + 3) * kGolden);
    const float r0 = (float)(int32_t)((t & 0x1fffffull) % 1001ull);
    const float r1 = (float)(int32_t)(((t >> 21) & 0x1fffffull) % 1001ull);
    const float r2 = (float)(int32_t)(((t >> 42) & 0x1fffffull) % 1001ull);
    const float kk = 0.1f * hfac * collision;
    const f3 tb = smul(kk, subs(mk(r0, r1, r2), 500.0f));
    sp = add(sp, mk(swe_div500(tb.x), swe_div500(tb.y), swe_div500(tb.z)));
  }

  sp = muls(sp, (float)(1.0 - 0.3 * (double)w.sed));  // :87 (double, then float)

  w.px += sp.x;  // :91
  w.py += sp.y;
  w.pz += sp.z;

  if (TRACK) acc.add_track(cell, w.sed, sp);  // :93-96 on the pre-move cell

  float force = -dot(normalize(sp), n) * length(sp);  // :100-102
  if (force < 0.0f) force = 0.0f;
  const float lift = (1.0f - collision) * length(sp);  // :104
  const float capacity = force * hfac + 0.02f * lift * hfac;  // :106

  const float diff = capacity - w.sed;  // :110-112
  float q;
  if (LAYERS == 2 && diff > 0.0f && map.s0() <= 0.0f) {
    q = k.abrasion * diff;  // bare bedrock: abrasion rate, sediment layer untouched
    const int64_t qf = fx_from_f32(q);
    acc.add_pickup(cell, -qf, false);
    w.ledger += qf;
  } else {
    q = k.suspension * diff;
    const int64_t qf = fx_from_f32(q);
    acc.add_pickup(cell, -qf, LAYERS == 2);
    w.ledger += qf;
  }
  w.sed += q;
  w.sx = sp.x;
  w.sy = sp.y;
  w.sz = sp.z;
  w.lastc = (int32_t)cell;
  acc.add_event(cell);  // :114 World::cascade(ipos), applied at the window's end
  return kStepped;
}

}  // namespace swe
