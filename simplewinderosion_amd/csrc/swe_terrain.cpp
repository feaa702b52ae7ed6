// swe_terrain.cpp -- seeded input terrains for the erosion workloads.
//
// The reference initialises its map with a Gaussian bump (cellpool.h:378-385,
// 416-428) and keeps an OpenSimplex2 fBm generator commented out
// (cellpool.h:365-367,388-411).  The benchmark workloads need a periodic fBm
// gradient-noise terrain with optional square pyramids; this file generates
// it on the host.  Recipe (the structure of FastNoiseLite's fBm Perlin,
// FastNoiseLite.h:844-862,1769-1792, with this project's own hash and
// gradient set): per octave a lattice of unit gradients chosen by a 64-bit
// hash of (seed, octave, lattice point), periodic across the map; quintic
// fade 6t^5-15t^4+10t^3; amplitude halves and frequency doubles per octave;
// the sum is min-max scaled to [0,1]; pyramids are added afterwards.
#include <cmath>
#include <algorithm>
#include <cstdint>
#include <vector>

#include "../../include/swe.h"

namespace {

uint64_t mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }

// gradient dot offset at lattice point (i, j) of octave o
double grad(uint64_t seed, int o, int64_t i, int64_t j, double dx, double dy) {
  const uint64_t h = mix(seed ^ mix(0x9E3779B97F4A7C15ull * (uint64_t)(o + 1) + mix((uint64_t)i * 0x632BE59BD9B4E019ull + (uint64_t)j)));
  const double ang = (double)(h >> 40) * (6.283185307179586 / 16777216.0);
  return dx * std::cos(ang) + dy * std::sin(ang);
}

double uniform(uint64_t& state) {
  state += 0x9E3779B97F4A7C15ull;
  return (double)(mix(state) >> 11) * (1.0 / 9007199254740992.0);
}

}  // namespace

extern "C" int swe_generate_terrain(int32_t nx, int32_t ny, uint64_t seed, int32_t octaves, float base_cycles,
                                    int32_t n_pyramids, float pyr_hmin, float pyr_hmax, float pyr_bmin,
                                    float pyr_bmax, float* out) {
  if (nx < 1 || ny < 1 || !out || octaves < 1 || octaves > 24 || !(base_cycles > 0.0f) || n_pyramids < 0)
    return SWE_EINVAL;
  const int64_t n = (int64_t)nx * ny;
  std::vector<double> h((size_t)n, 0.0);
  double amp = 1.0;
  for (int o = 0; o < octaves; o++) {
    const double f = (double)base_cycles * std::ldexp(1.0, o);  // periods across x
    const int64_t px = std::max<int64_t>(1, (int64_t)std::llround(f));
    const int64_t py = std::max<int64_t>(1, (int64_t)std::llround(f * (double)ny / (double)nx));
    for (int32_t x = 0; x < nx; x++) {
      const double u = (double)x * (double)px / (double)nx;
      const int64_t i0 = (int64_t)std::floor(u);
      const double dx = u - (double)i0;
      const double sx = fade(dx);
      const int64_t ia = i0 % px, ib = (i0 + 1) % px;
      for (int32_t y = 0; y < ny; y++) {
        const double v = (double)y * (double)py / (double)ny;
        const int64_t j0 = (int64_t)std::floor(v);
        const double dy = v - (double)j0;
        const double sy = fade(dy);
        const int64_t ja = j0 % py, jb = (j0 + 1) % py;
        const double g00 = grad(seed, o, ia, ja, dx, dy);
        const double g10 = grad(seed, o, ib, ja, dx - 1.0, dy);
        const double g01 = grad(seed, o, ia, jb, dx, dy - 1.0);
        const double g11 = grad(seed, o, ib, jb, dx - 1.0, dy - 1.0);
        const double a = g00 + sx * (g10 - g00);
        const double b = g01 + sx * (g11 - g01);
        h[(size_t)x * ny + y] += amp * (a + sy * (b - a));
      }
    }
    amp *= 0.5;
  }
  double lo = h[0], hi = h[0];
  for (double v : h) {
    lo = v < lo ? v : lo;
    hi = v > hi ? v : hi;
  }
  const double span = hi > lo ? hi - lo : 1.0;
  for (int64_t c = 0; c < n; c++) out[c] = (float)((h[(size_t)c] - lo) / span);

  // square pyramids (height layer only), periodic distance
  uint64_t st = mix(seed ^ 0x5059524D49445321ull);
  for (int32_t k = 0; k < n_pyramids; k++) {
    const double cx = uniform(st) * nx, cy = uniform(st) * ny;
    const double base = pyr_bmin + uniform(st) * (pyr_bmax - pyr_bmin);
    const double top = pyr_hmin + uniform(st) * (pyr_hmax - pyr_hmin);
    const double half = 0.5 * base;
    if (!(half > 0.0)) continue;
    const int32_t r = (int32_t)std::ceil(half);
    for (int32_t dx = -r; dx <= r; dx++)
      for (int32_t dy = -r; dy <= r; dy++) {
        const int64_t gx = (int64_t)std::floor(cx) + dx, gy = (int64_t)std::floor(cy) + dy;
        const double ddx = std::fabs((double)gx + 0.5 - cx), ddy = std::fabs((double)gy + 0.5 - cy);
        const double m = ddx > ddy ? ddx : ddy;
        if (m >= half) continue;
        const int64_t wx = ((gx % nx) + nx) % nx, wy = ((gy % ny) + ny) % ny;
        out[wx * ny + wy] += (float)(top * (1.0 - m / half));
      }
  }
  return 0;
}
