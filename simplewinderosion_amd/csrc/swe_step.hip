// swe_step.hip -- spawn kernels of the tile-owned step path and the
// dispatch of a window's step kernel to the compiled tile shape.
#include "swe_fly.hpp"
#include "swe_kernels.h"
#include "swe_step.h"

namespace swe {

// ---------------------------------------------------------------------------
// Spawn and bin the particles of a cycle by tile (world.hpp:52-62)
// ---------------------------------------------------------------------------
// terrain of a global particle index: last k with part0 <= i
__device__ __forceinline__ int32_t terrain_of(const Terrain* terr, int32_t nterr, int64_t i) {
  int32_t lo = 0, hi = nterr - 1;
  while (lo < hi) {
    const int32_t mid = (lo + hi + 1) >> 1;
    if (terr[mid].part0 <= i) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_spawn_count(SpawnArgs a) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= a.n) return;
  const Terrain T = a.terr[terrain_of(a.terr, a.nterr, i)];
  Particle w;
  spawn(particle_key(cycle_key(T.seed, *a.cycle), (uint64_t)(i - T.part0)), T.nx, T.ny, a.spawn_mode, w);
  // the spawn cell is always on the map: round() of a position in
  // [ix, ix + 0.999] is ix or ix + 1, wrapped (periodic) or, on a bounded map,
  // possibly nx / ny, in which case the particle dies at its first call;
  // bin it by the spawn cell clamped to the map
  int32_t ix = (int32_t)__builtin_roundf(w.px), iy = (int32_t)__builtin_roundf(w.pz);
  if (a.periodic) {
    if (ix >= T.nx) ix -= T.nx;
    if (iy >= T.ny) iy -= T.ny;
  } else {
    if (ix >= T.nx) ix = T.nx - 1;
    if (iy >= T.ny) iy = T.ny - 1;
  }
  const int32_t t = T.tile0 + (ix / a.tile_sx) * T.tiles_y + (iy / a.tile_sy);
  a.tile[i] = t;
  a.rank[i] = atomicAdd(&a.count[t], 1u);  // count is window 0's nin
}

__global__ __launch_bounds__(256) void k_spawn_scan(SpawnArgs a) {
  // one workgroup: exclusive prefix of the per-tile counts -> each tile's
  // spawned particles form sub-range 4 (its own tile) of window 0's input;
  // chunk sums of the counts
  __shared__ int64_t s_part[256];
  const int32_t ntiles = a.ntiles;
  const int32_t per = (ntiles + 255) / 256;
  const int32_t t0 = threadIdx.x * per, t1 = t0 + per < ntiles ? t0 + per : ntiles;
  int64_t run = 0;
  for (int32_t t = t0; t < t1; t++) run += a.count[t];
  s_part[threadIdx.x] = run;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t acc = 0;
    for (int i = 0; i < 256; i++) {
      const int64_t v = s_part[i];
      s_part[i] = acc;
      acc += v;
    }
  }
  __syncthreads();
  int64_t b = s_part[threadIdx.x];
  for (int32_t t = t0; t < t1; t++) {
    const uint32_t n = a.count[t];
    for (int d = 0; d < 9; d++) {
      a.sub_base[(int64_t)t * 9 + d] = d == 4 ? b : 0;
      a.sub_cnt[(int64_t)t * 9 + d] = d == 4 ? n : 0u;
    }
    b += n;
  }
  const int32_t nchunks = (ntiles + kStepChunk - 1) / kStepChunk;
  for (int32_t c = threadIdx.x; c < nchunks; c += 256) {
    uint32_t sum = 0;
    const int32_t e = (c + 1) * kStepChunk < ntiles ? (c + 1) * kStepChunk : ntiles;
    for (int32_t t = c * kStepChunk; t < e; t++) sum += a.count[t];
    a.csum[c] = sum;
  }
}

__global__ __launch_bounds__(256) void k_spawn_place(SpawnArgs a) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= a.n) return;
  const Terrain T = a.terr[terrain_of(a.terr, a.nterr, i)];
  Particle w;
  spawn(particle_key(cycle_key(T.seed, *a.cycle), (uint64_t)(i - T.part0)), T.nx, T.ny, a.spawn_mode, w);
  const int64_t o = a.sub_base[(int64_t)a.tile[i] * 9 + 4] + a.rank[i];
  a.out[o] = PRec{f4a8{w.px, w.py, w.pz, w.sed}, f4a8{w.sx, w.sy, w.sz, u2f((uint32_t)i)}, 0};
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
void launch_step(int shape, const StepArgs& a, bool periodic, bool track, int layers, bool deposit, int mode,
                 hipStream_t s) {
  switch (shape) {
#define SWE_SHAPE_CASE(i)                                                           \
  case i:                                                                           \
    launch_step_shape##i(a, periodic, track, layers, deposit, mode, s); \
    break;
    SWE_SHAPE_CASE(0)
    SWE_SHAPE_CASE(1)
    SWE_SHAPE_CASE(2)
    SWE_SHAPE_CASE(3)
    SWE_SHAPE_CASE(4)
#undef SWE_SHAPE_CASE
    default:
      break;
  }
}

void launch_spawn(const SpawnArgs& a, hipStream_t s) {
  if (a.n > 0) {
    const unsigned g = (unsigned)((a.n + 255) / 256);
    hipLaunchKernelGGL(k_spawn_count, dim3(g), dim3(256), 0, s, a);
  }
  hipLaunchKernelGGL(k_spawn_scan, dim3(1), dim3(256), 0, s, a);
  if (a.n > 0) {
    const unsigned g = (unsigned)((a.n + 255) / 256);
    hipLaunchKernelGGL(k_spawn_place, dim3(g), dim3(256), 0, s, a);
  }
}

}  // namespace swe
