// swe_cascade.hpp -- one cell's cascade of a snapshot window on the device.
#pragma once

#include "swe_device.hpp"

namespace swe {

// Settling factor of a cell's `events` cascade events in one window:
// settling for one event, else 1 - (1 - settling)^events by binary
// exponentiation (least significant bit first) in float.
__device__ __forceinline__ float settle_factor_dev(float settling, uint32_t events) {
  if (events <= 1u) return settling;
  float b = 1.0f - settling, r = 1.0f;
  uint32_t e = events;
  while (e) {
    if (e & 1u) r = r * b;
    b = b * b;
    e >>= 1;
  }
  return 1.0f - r;
}

// Neighbour offsets of World::cascade (world.hpp:83-92), k = 0..7:
// (-1,-1) (-1,0) (-1,1) (0,-1) (0,1) (1,-1) (1,0) (1,1), packed two bits per
// k as offset + 1; bit k of kDiagMask marks the diagonals (distance sqrt 2).
constexpr uint32_t kOffXPacked = 0xa940u, kOffYPacked = 0x9224u, kDiagMask = 0xa5u;
__device__ __forceinline__ int32_t off_x(int32_t k) { return (int32_t)((kOffXPacked >> (2 * k)) & 3u) - 1; }
__device__ __forceinline__ int32_t off_y(int32_t k) { return (int32_t)((kOffYPacked >> (2 * k)) & 3u) - 1; }

// Sort key of a neighbour: descending order of the key is descending height,
// ties in ascending offset order -- the order the stable insertion sort of
// world.hpp:118-120 leaves eight elements in.  The height bits are mapped to
// an unsigned integer of the same order (-0 first folded into +0, which the
// float comparison treats as equal); map heights are finite.  Out-of-map
// neighbours get key 0, below every in-map key.
__device__ __forceinline__ uint64_t cascade_key(float h, int32_t k, bool valid) {
  const uint32_t u = f2u(h + 0.0f);
  const uint32_t o = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  return valid ? (((uint64_t)o << 32) | (uint32_t)(7 - k)) : 0ull;
}
__device__ __forceinline__ float cascade_key_height(uint64_t key) {
  const uint32_t o = (uint32_t)(key >> 32);
  return u2f((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

// One cell's cascade of the window: World::cascade (world.hpp:79-153)
// against the snapshot with the window's settling factor.  hn[k]: neighbour
// k's height in the fixed offset order; sed_of(k): its sediment (two-layer
// maps only); hc / sc: the centre's; bit k of validmask set for in-map
// neighbours.  The neighbours are sorted by a sorting network on their keys
// and walked in that order with the centre evolving; emit(k, amount) is
// called with the fixed-point amount neighbour k receives (the centre loses
// the same amount) for every transfer.
template <int LAYERS, class SedOf, class Emit>
__device__ __forceinline__ void cascade_cell_dev(const float hn[8], const SedOf& sed_of, float hc, float sc,
                                                 uint32_t validmask, float settling, float maxdiff, Emit&& emit) {
  uint64_t key[8];
#pragma unroll
  for (int k = 0; k < 8; k++) key[k] = cascade_key(hn[k], k, (validmask >> k) & 1u);
#define SWE_CAS(i, j)                        \
  {                                          \
    const uint64_t a = key[i], b = key[j];   \
    const bool sw = b > a;                   \
    key[i] = sw ? b : a;                     \
    key[j] = sw ? a : b;                     \
  }
  SWE_CAS(0, 2) SWE_CAS(1, 3) SWE_CAS(4, 6) SWE_CAS(5, 7)
  SWE_CAS(0, 4) SWE_CAS(1, 5) SWE_CAS(2, 6) SWE_CAS(3, 7)
  SWE_CAS(0, 1) SWE_CAS(2, 3) SWE_CAS(4, 5) SWE_CAS(6, 7)
  SWE_CAS(2, 4) SWE_CAS(3, 5)
  SWE_CAS(1, 4) SWE_CAS(3, 6)
  SWE_CAS(1, 2) SWE_CAS(3, 4) SWE_CAS(5, 6)
#undef SWE_CAS
  const float dd = __builtin_sqrtf(2.0f);
#pragma unroll
  for (int r = 0; r < 8; r++) {
    if (key[r] == 0ull) break;  // the out-of-map neighbours come last
    const int32_t k = 7 - (int32_t)(key[r] & 7u);
    const float diff = hc - cascade_key_height(key[r]);
    const float d = ((kDiagMask >> k) & 1u) ? dd : 1.0f;
    const float excess = __builtin_fabsf(diff) - (d * maxdiff) * 1.0f;
    float t = settling * excess / 2.0f;
    bool go = diff != 0.0f && !(excess <= 0.0f);
    if (LAYERS == 2 && go) {
      const float avail = diff > 0.0f ? sc : sed_of(k);
      if (t > avail) t = avail;
      go = t > 0.0f;
    }
    if (!go) continue;
    const int64_t tf = fx_from_f32(t);
    if (diff > 0.0f) {
      hc -= t;
      sc -= t;
      emit(k, tf);
    } else {
      hc += t;
      sc += t;
      emit(k, -tf);
    }
  }
}

}  // namespace swe
