// swe_step.h -- argument blocks of the tile-owned step kernel (swe_step.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace swe {

// Tile shapes of the step kernel: SX x SY cells (x the slow axis), NT
// threads, a register budget that keeps WPE waves per SIMD.  Every shape
// computes the same bits; swe_autotune times them (SWE_TUNE_SHAPE).
struct StepShape {
  int32_t sx, sy, threads, waves;
};
constexpr int kStepShapes = 5;
constexpr StepShape kStepShape[kStepShapes] = {
    {16, 32, 256, 5},  // default: the best measured on the 1024^2 and 64-terrain workloads
    {32, 16, 256, 5},
    {16, 16, 128, 5},
    {16, 32, 256, 4},
    {32, 32, 512, 4},
};
constexpr int kStepChunk = 32;                 // tiles per chunk sum of the input counts

// One terrain of a context: its cells, tiles and particles are contiguous
// ranges of the context's buffers
struct Terrain {
  int64_t cell0;   // first cell in the map arrays
  int64_t part0;   // first particle index of a cycle
  int64_t npart;   // particles per cycle
  int32_t nx, ny, tiles_x, tiles_y, tile0;
  float rcp_nx, rcp_ny;
  uint64_t seed;
  float ps[3], nps[3];  // pspeed and normalize(pspeed)
};

// One particle record: 40 bytes (two 16-byte words and the ledger), moved
// with three vector loads or stores at immediate offsets from one address
// (8-byte aligned; global loads and stores of 16 bytes need only 4).  The
// number of calls a particle has made (its age) is the window index: every
// live particle makes one call per window from window 0.  A particle whose
// position left a bounded map (or stopped being finite) is stored with the
// bits of its last cell in q0.x and a NaN in q0.z: its next call finds no
// cell, and it dies (and deposits) at that cell.
typedef float f4a8 __attribute__((ext_vector_type(4), aligned(8)));
struct PRec {
  f4a8 q0;         // pos.xyz, sediment (off the map: last cell's bits, -, NaN, sediment)
  f4a8 q1;         // speed.xyz, particle index bits (RNG stream)
  int64_t ledger;  // fixed-point load
};
static_assert(sizeof(PRec) == 40 && alignof(PRec) == 8, "particle record: 40 bytes");

struct StepArgs {
  // map, window parity in -> out.  A master holds the previous window's
  // pickups already; its cascades are added here (relax), then this
  // window's pickups before it is written out.
  const int64_t* zm_in;
  int64_t* zm_out;
  const float* snap_in;  // snapshot the previous window's particles read
  float* snap_out;       // snapshot of this window (master after the relax)
  const int64_t* sm_in;
  int64_t* sm_out;
  const float* ssnap_in;
  float* ssnap_out;
  const uint32_t* v_in;  // previous window's cascade events
  uint32_t* v_out;
  int64_t* trk;          // 4 x ncell tracks of the cycle
  int64_t* lost;         // kLostSlots counters
  uint64_t* stat;        // kStatKinds x kStatSlots statistics
  // particles
  const PRec* in;
  PRec* out;
  const PRec* far_in;
  PRec* far_out;
  uint32_t* far_cnt_in;
  uint32_t* far_cnt_out;
  uint32_t* far_cnt_clear;   // the far counter of the window after next
  // routing: 9 sub-ranges (one per direction) of each tile's output region
  const int64_t* sub_base_prev;  // [tile * 9 + d] start of sub-range d in `in`
  const uint32_t* sub_cnt_prev;  // [tile * 9 + d] its length
  int64_t* sub_base_cur;         // the same for `out` (written here)
  uint32_t* sub_cnt_cur;
  const uint32_t* nin_cur;   // particles each tile receives this window
  uint32_t* nin_next;        // ... next window (accumulated here)
  uint32_t* nin_clear;       // ... the window after next (zeroed here)
  const uint32_t* csum_cur;  // nin_cur summed over chunks of kStepChunk tiles
  uint32_t* csum_next;
  uint32_t* csum_clear;
  const uint64_t* cycle;
  const Terrain* terr;
  const int32_t* tile_terrain;  // per global tile
  int32_t nterr;                // 1: every tile is terrain 0 (no tile_terrain lookup)
  int64_t ncell;                // cells of all terrains (track plane stride)
  int32_t ntiles;
  float mapscale, boundarylayer, suspension, gravity, abrasion, maxdiff, settling;
  int32_t max_age;
  // workgroup -> tile order: 0 = each XCD a contiguous range of tiles;
  // C >= 1 = runs of C consecutive tiles dealt round-robin over the XCDs
  int32_t tile_order;
  // crowded tiles first: a tile with more than kHeavyMin arrivals in one
  // window is listed for the next launch, whose first heavy_slots
  // workgroups take the listed tiles (so they start at once instead of
  // finishing last); the workgroup of the usual order then skips the tile
  int32_t heavy_slots;       // 0 = off, else a multiple of 8 (XCD mapping)
  int32_t default_slots;     // workgroups of the usual order: ntiles (a multiple of 8 with light slots)
  int32_t light_slots;       // 0 = off, else a multiple of 8: light tiles last, per XCD
  int32_t window;            // window index within the cycle
  const uint32_t* heavy_cur; // this launch's listed tiles: kHeavySlotsMax heavy, then light_slots light,
                             // each part in 8 equal per-XCD lists
  uint32_t* heavy_next;      // the next launch's (appended here)
  const uint32_t* hcnt_cur;  // list lengths (may exceed the slots): [x] heavy, [8 + x] light on XCD x
  uint32_t* hcnt_next;
  uint32_t* hcnt_clear;      // the lengths of the lists after next
  uint32_t* hstamp;          // per tile: w + 1 when listed for window w (0 at cycle start)
};

#ifndef SWE_HEAVY_MIN
#define SWE_HEAVY_MIN 1024
#endif
#ifndef SWE_HEAVY_SLOTS
#define SWE_HEAVY_SLOTS 256
#endif
#ifndef SWE_LIGHT_MAX
#define SWE_LIGHT_MAX 256
#endif
constexpr uint32_t kHeavyMin = SWE_HEAVY_MIN;        // arrivals that make a tile heavy
constexpr uint32_t kLightMax = SWE_LIGHT_MAX;        // arrivals that leave a tile light
constexpr int32_t kHeavySlotsMax = SWE_HEAVY_SLOTS;  // listed tiles per launch (a multiple of 8)

struct SpawnArgs {
  PRec* out;
  int32_t* tile;     // per particle
  uint32_t* rank;    // per particle
  uint32_t* count;   // per tile, zero on entry: window 0's input counts
  uint32_t* csum;    // per chunk of kStepChunk tiles: window 0's chunk sums
  int64_t* sub_base; // [tile * 9 + d]: window 0's arrivals, all in d = 4 (own tile)
  uint32_t* sub_cnt;
  const uint64_t* cycle;
  const Terrain* terr;
  int32_t nterr;
  int64_t n;         // particles of all terrains
  int32_t ntiles;
  int32_t periodic, spawn_mode;
  int32_t tile_sx, tile_sy;  // the tile shape the particles are binned for
};

void launch_step(int shape, const StepArgs& a, bool periodic, bool track, int layers, bool deposit,
                 int mode, hipStream_t s);
#define SWE_DECLARE_SHAPE(i)                                                                                    \
  void launch_step_shape##i(const StepArgs& a, bool periodic, bool track, int layers, bool deposit, int mode, \
                            hipStream_t s);
SWE_DECLARE_SHAPE(0)
SWE_DECLARE_SHAPE(1)
SWE_DECLARE_SHAPE(2)
SWE_DECLARE_SHAPE(3)
SWE_DECLARE_SHAPE(4)
#undef SWE_DECLARE_SHAPE
void launch_spawn(const SpawnArgs& a, hipStream_t s);

}  // namespace swe
