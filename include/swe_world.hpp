// swe_world.hpp -- C++ mirror of SimpleWindErosion's simulation interface
// over libswe.so (C-ABI in swe.h).
//
// Replaces, for a C++ caller of the reference:
//   struct Wind statics            source/wind.hpp:13,15-26
//   class World statics + erode    source/world.hpp:12-32,36-77
//   World::map + map.init          source/world.hpp:13,28; source/cellpool.h:339-430
// As in the reference the parameters are static members read at every
// World::erode(cycles) call; the map lives on the GPU and is read back with
// World::download().  Header-only; link with -lswe.
#ifndef SWE_WORLD_HPP
#define SWE_WORLD_HPP

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "swe.h"

namespace swe {

struct Wind {
  inline static int maxAge = 512;                  // wind.hpp:23
  inline static float boundarylayer = 2.0f;        // wind.hpp:24
  inline static float suspension = 0.05f;          // wind.hpp:25
  inline static float gravity = 0.1f;              // wind.hpp:26
  inline static float pspeed[3] = {1.0f, 0.0f, 0.0f};  // wind.hpp:13 normalize(vec3(1,0,0))
};

class World {
 public:
  inline static unsigned int SEED = 0;     // world.hpp:26
  inline static float lrate = 0.1f;        // world.hpp:30
  inline static float maxdiff = 0.005f;    // world.hpp:31
  inline static float settling = 1.0f;     // world.hpp:32

  // extensions of the reference's compile-time map (cellpool.h:184-198)
  inline static int nx = 512, ny = 512;
  inline static float mapscale = 80.0f;
  inline static int periodic = 0, track = 1, layers = 1, window = 1, deposit_on_death = 0, spawn = 0;
  inline static float abrasion = 0.0f;
  // -1 automatic: the colour-ordered cascade when settling > 0.25 (the
  // reference's 1.0 included), else the synchronous one
  inline static int cascade_order = -1;
  inline static int device = 0;

  // map.init(...) counterpart: create the device map and upload heights
  // (nx*ny floats, index x*ny + y as math::flatten) and optional sediment.
  static void init(const float* height, const float* sediment = nullptr) {
    release();
    check(swe_create(&ctx_, nx, ny, device), "swe_create");
    push();
    check(swe_upload_layers(ctx_, height, sediment), "swe_upload_layers");
  }

  // World::erode(int cycles): `cycles` particles, one map-cycle
  static void erode(int cycles) {
    if (!ctx_) throw std::runtime_error("World::init was not called");
    push();
    check(swe_erode(ctx_, cycles, 1), "swe_erode");
  }

  // heights after the last erode (synchronizes); fields optional
  static void download(std::vector<float>& height, std::vector<float>* momentum = nullptr,
                       std::vector<float>* massflow = nullptr) {
    height.resize((size_t)nx * ny);
    if (momentum) momentum->resize((size_t)3 * nx * ny);
    if (massflow) massflow->resize((size_t)nx * ny);
    check(swe_download_layers(ctx_, height.data(), nullptr, massflow ? massflow->data() : nullptr,
                              momentum ? momentum->data() : nullptr),
          "swe_download_layers");
  }

  // quad::map::normal for every cell (cellpool.h:200-223,478-480)
  static void normals(std::vector<float>& out) {
    out.resize((size_t)3 * nx * ny);
    check(swe_normals(ctx_, out.data()), "swe_normals");
  }

  static swe_ctx* context() { return ctx_; }

  static void release() {
    if (ctx_) swe_destroy(ctx_);
    ctx_ = nullptr;
  }

 private:
  inline static swe_ctx* ctx_ = nullptr;

  static void check(int rc, const char* what) {
    if (rc != 0) {
      std::string msg = std::string(what) + " failed (" + std::to_string(rc) + ")";
      if (ctx_) msg += ": " + std::string(swe_last_error(ctx_));
      throw std::runtime_error(msg);
    }
  }

  static void push() {
    swe_params p;
    swe_params_default(&p);
    p.max_age = Wind::maxAge;
    p.boundarylayer = Wind::boundarylayer;
    p.suspension = Wind::suspension;
    p.gravity = Wind::gravity;
    for (int k = 0; k < 3; k++) p.pspeed[k] = Wind::pspeed[k];
    p.lrate = lrate;
    p.maxdiff = maxdiff;
    p.settling = settling;
    p.abrasion = abrasion;
    p.mapscale = mapscale;
    p.window = window;
    p.periodic = periodic;
    p.track = track;
    p.layers = layers;
    p.deposit_on_death = deposit_on_death;
    p.spawn = spawn;
    p.seed = SEED;
    p.cascade_order = cascade_order;
    check(swe_set_params(ctx_, &p), "swe_set_params");
  }
};

}  // namespace swe

#endif  // SWE_WORLD_HPP
