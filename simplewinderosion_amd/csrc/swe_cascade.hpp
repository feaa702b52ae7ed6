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

// One cell's cascade of the window: World::cascade (world.hpp:79-153)
// against the snapshot with the window's settling factor.  hn[k] / sn[k]:
// neighbour k's height / sediment in the fixed offset order
// (world.hpp:83-92); hc / sc: the centre's; bit k of validmask set for
// in-map neighbours.  The neighbours are sorted by height descending, ties
// in offset order (what the stable insertion sort std::sort uses on eight
// elements does), by a sorting network on (height, offset) keys; the walk
// then evolves the centre.  amt[k] = fixed-point amount neighbour k receives.
template <int LAYERS>
__device__ __forceinline__ void cascade_cell_dev(const float hn[8], const float sn[8], float hc, float sc,
                                             uint32_t validmask, float settling, float maxdiff, int64_t amt[8]) {
  float key[8], sed[8];
  int32_t idx[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    key[k] = ((validmask >> k) & 1u) ? hn[k] : -__builtin_inff();
    sed[k] = LAYERS == 2 ? sn[k] : 0.0f;
    idx[k] = k;
  }
#define SWE_CAS(i, j)                                                                  \
  {                                                                                    \
    const bool sw = key[j] > key[i] || (key[j] == key[i] && idx[j] < idx[i]);          \
    const float tk = sw ? key[j] : key[i];                                             \
    const int32_t ti = sw ? idx[j] : idx[i];                                           \
    key[j] = sw ? key[i] : key[j];                                                     \
    idx[j] = sw ? idx[i] : idx[j];                                                     \
    key[i] = tk;                                                                       \
    idx[i] = ti;                                                                       \
    if (LAYERS == 2) {                                                                 \
      const float ts = sw ? sed[j] : sed[i];                                           \
      sed[j] = sw ? sed[i] : sed[j];                                                   \
      sed[i] = ts;                                                                     \
    }                                                                                  \
  }
  SWE_CAS(0, 2) SWE_CAS(1, 3) SWE_CAS(4, 6) SWE_CAS(5, 7)
  SWE_CAS(0, 4) SWE_CAS(1, 5) SWE_CAS(2, 6) SWE_CAS(3, 7)
  SWE_CAS(0, 1) SWE_CAS(2, 3) SWE_CAS(4, 5) SWE_CAS(6, 7)
  SWE_CAS(2, 4) SWE_CAS(3, 5)
  SWE_CAS(1, 4) SWE_CAS(3, 6)
  SWE_CAS(1, 2) SWE_CAS(3, 4) SWE_CAS(5, 6)
#undef SWE_CAS
  const float dd = __builtin_sqrtf(2.0f);
  int64_t got[8];  // in sorted slot order
#pragma unroll
  for (int r = 0; r < 8; r++) {
    got[r] = 0;
    const int32_t k = idx[r];
    if (!((validmask >> k) & 1u)) continue;
    const float diff = hc - key[r];
    if (diff == 0.0f) continue;
    const float d = (k == 0 || k == 2 || k == 5 || k == 7) ? dd : 1.0f;
    const float excess = __builtin_fabsf(diff) - (d * maxdiff) * 1.0f;
    if (excess <= 0.0f) continue;
    float t = settling * excess / 2.0f;
    if (LAYERS == 2) {
      const float avail = diff > 0.0f ? sc : sed[r];
      if (t > avail) t = avail;
      if (!(t > 0.0f)) continue;
    }
    const int64_t tf = fx_from_f32(t);
    if (diff > 0.0f) {
      hc -= t;
      sc -= t;
      got[r] = tf;
    } else {
      hc += t;
      sc += t;
      got[r] = -tf;
    }
  }
  // back to offset order
#pragma unroll
  for (int k = 0; k < 8; k++) {
    int64_t v = 0;
#pragma unroll
    for (int r = 0; r < 8; r++) v = idx[r] == k ? got[r] : v;
    amt[k] = v;
  }
}

}  // namespace swe
