// erode_demo.cpp -- the reference's main loop reduced to its erosion calls,
// written against include/swe_world.hpp: World::init, then World::erode(N)
// per frame (SimpleWindErosion.cpp:36-38,212-217).  Prints the exact mass
// sum and a height checksum so that tests can compare with other callers.
//   erode_demo NX NY SEED PARTICLES FRAMES [reference]
// With "reference" every static keeps the reference's value (bounded map,
// settling 1.0, suspension 0.05, ...); otherwise the map is periodic and
// settling 0.05.
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "swe_world.hpp"

int main(int argc, char** argv) {
  if (argc < 6) {
    std::fprintf(stderr, "usage: %s NX NY SEED PARTICLES FRAMES [reference]\n", argv[0]);
    return 2;
  }
  const bool reference = argc > 6 && std::string(argv[6]) == "reference";
  const int nx = std::atoi(argv[1]), ny = std::atoi(argv[2]);
  const unsigned seed = (unsigned)std::strtoul(argv[3], nullptr, 10);
  const int particles = std::atoi(argv[4]), frames = std::atoi(argv[5]);
  std::vector<float> h((size_t)nx * ny);
  if (swe_generate_terrain(nx, ny, seed, 6, 4.0f, 0, 0, 0, 0, 0, h.data()) != 0) return 1;
  for (float& v : h) v *= 0.5f;
  swe::World::nx = nx;
  swe::World::ny = ny;
  swe::World::SEED = seed;
  if (!reference) {
    swe::World::settling = 0.05f;
    swe::World::periodic = 1;
  }
  try {
    swe::World::init(h.data());
    for (int f = 0; f < frames; f++) swe::World::erode(particles);
    std::vector<float> out;
    swe::World::download(out);
    swe_audit_t a;
    if (swe_audit(swe::World::context(), &a) != 0) return 1;
    double cs = 0.0;
    for (float v : out) cs += v;
    std::printf("{\"height_sum_lo\": %" PRId64 ", \"height_sum_hi\": %" PRId64 ", \"lost\": %" PRId64
                ", \"checksum\": %.9g}\n",
                a.height_sum_lo, a.height_sum_hi, a.lost, cs);
    swe::World::release();
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
