// swe_kernels.h -- argument blocks and launchers of the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/swe.h"

namespace swe {

// carried-off mass is spread over this many counters to avoid one hot address
constexpr int kLostSlots = 256;
// device statistics: kStatKinds counters of kStatSlots slots each (spread
// over slots to keep the atomics off one address); summed (max for
// kStatMaxTile) by the host
constexpr int kStatSlots = 64;
enum StatKind { kStatSteps = 0, kStatMaxTile = 1, kStatCrowded = 2, kStatFar = 3, kStatHeavy = 4, kStatLight = 5, kStatKinds = 6 };

struct MarchArgs {
  const float* S;      // height snapshot of the window
  const float* SS;     // sediment snapshot (layers == 2)
  int64_t* D;          // height pickup deltas
  int64_t* DS;         // sediment pickup deltas (layers == 2)
  uint32_t* V;         // cascade events per cell
  int64_t* T;          // 4 x ncell tracks (track)
  float4* P0;          // pos.xyz, sediment
  float4* P1;          // speed.xyz, age bits
  int64_t* L;          // fixed-point load
  int32_t* LC;         // last pre-move cell
  int64_t* lost;       // kLostSlots counters
  uint64_t* stat;      // kStatKinds x kStatSlots
  const uint64_t* cycle;
  int64_t n;           // particles
  int64_t ncell;
  int32_t nx, ny;
  float rcp_nx, rcp_ny;
  int32_t R;           // fly() calls in this window
  float mapscale;
  float ps[3];         // pspeed
  float nps[3];        // normalize(pspeed)
  swe_params p;
};

struct RelaxArgs {
  const float* S;      // window snapshot (read, with halo)
  const float* SS;
  const uint32_t* V;   // window event counts (read, with halo)
  uint32_t* v_clear;   // the other event buffer, zeroed for the next window
  const int64_t* zm_in;  // master height of the window
  int64_t* zm_out;       // master height after the window
  int64_t* dz;           // pickup deltas (read, then zeroed)
  float* zs_out;         // next snapshot
  const int64_t* sm_in;
  int64_t* sm_out;
  int64_t* ds;
  float* ss_out;
  int32_t nx, ny;
  float rcp_nx, rcp_ny;
  float maxdiff, settling;
};

// Colour classes of the colour-ordered cascade along one axis of n cells.
// Bounded: x mod 3.  Periodic, n = 3q + r: x mod 3 for x < 3q and each of
// the last r cells a class of its own, so that two cells of one class are at
// least three apart across the wrap too (their 3x3 neighbourhoods are
// disjoint).  Classes run in order, x class major.
__host__ __device__ inline int32_t colour_classes(int32_t n, bool periodic) { return periodic ? 3 + n % 3 : 3; }
__host__ __device__ inline int32_t colour_count(int32_t n, bool periodic, int32_t col) {
  const int32_t lim = periodic ? n - n % 3 : n;
  if (col < 3) return lim > col ? (lim - col + 2) / 3 : 0;
  return (periodic && col - 3 < n % 3) ? 1 : 0;
}
__host__ __device__ inline int32_t colour_cell(int32_t n, bool periodic, int32_t col, int32_t i) {
  (void)periodic;
  return col < 3 ? col + 3 * i : n - n % 3 + (col - 3);
}
constexpr int32_t kColourAxisMax = 5;  // classes along an axis: 3 + n % 3 at most
// Settling factors above this use the colour-ordered cascade when
// cascade_order is automatic; the synchronous cascade rejects them.
constexpr float kSettlingMaxSync = 0.25f;

void launch_march(const MarchArgs& a, bool first, hipStream_t s);
void launch_relax(const RelaxArgs& a, bool periodic, int layers, hipStream_t s);
// the window's colour-ordered relax: with `begin`, masters += pickups and
// their snapshots first (the global march leaves the pickups in dz / ds);
// then one launch per colour class on zm_out / zs_out (sm_out / ss_out) with
// the events V, and the sediment layer clamped at zero
void launch_relax_colour(const RelaxArgs& a, bool periodic, int layers, bool begin, hipStream_t s);
// the same relax over all terrains of a batch (masters already hold the
// window's pickups): class (cx, cy) of every terrain in one launch, classes
// x-major over a 5 x 5 grid (each terrain sees its own classes in its own
// order), then the sediment clamp over all ncell cells.  pre: the
// colour_prefix table on the device and on the host.
struct Terrain;
void launch_relax_colour_batch(const RelaxArgs& a, const Terrain* terr, const int64_t* pre_dev,
                               const int64_t* pre_host, int32_t nterr, int64_t ncell, bool periodic, int layers,
                               hipStream_t s);
// pre[(cx * 5 + cy) * (nterr + 1) + k]: cells of class (cx, cy) on terrains 0 .. k-1
void colour_prefix(const int32_t* nx, const int32_t* ny, int32_t nterr, bool periodic, int64_t* pre);
void launch_zero64(int64_t* p, int64_t n, hipStream_t s);
void launch_zero32(uint32_t* p, int64_t n, hipStream_t s);
void launch_from_f32(const float* in, int64_t* master, float* snap, int64_t n, hipStream_t s);
void launch_snap(const int64_t* master, float* snap, int64_t n, hipStream_t s);
void launch_ema(float* ema, const int64_t* trk, int64_t n4, float lrate, hipStream_t s);
void launch_bump(uint64_t* cycle, hipStream_t s);
void launch_windmap(const float* mom, int64_t plane, float* out, int64_t n, hipStream_t s);
void launch_mesh(const float* S, float* out, int32_t nx, int32_t ny, float mapscale, bool periodic, hipStream_t s);
void launch_normals(const float* S, float* out, int32_t nx, int32_t ny, float mapscale, bool periodic,
                    hipStream_t s);

// Known-answer entry points (not on the hot path): one fly() call of each
// of n caller-given particles, and one cell's cascade, against a snapshot,
// with the device functions the step kernels run.
struct DebugFlyArgs {
  const float* S;       // height snapshot
  const float* SS;      // sediment snapshot (layers == 2)
  float* st;            // n x 7: pos xyz, speed xyz, sediment (in/out)
  int32_t* age;         // n, in/out
  const uint64_t* key;  // n particle RNG keys
  int32_t* alive;       // n out: 1 stepped, 0 died
  int64_t* pick;        // n out: fixed-point amount taken from the cell
  int32_t n, nx, ny;
  float rcp_nx, rcp_ny, mapscale;
  float ps[3], nps[3];
  swe_params p;
};
void launch_debug_fly(const DebugFlyArgs& a, hipStream_t s);

struct DebugCascadeArgs {
  const float* S;
  const float* SS;
  int32_t nx, ny, x, y;
  uint32_t events;
  int32_t* count;  // out: in-map neighbours
  int64_t* cell;   // out [8]: their cells in offset order
  int64_t* fx;     // out [8]: the fixed-point amount each receives
  swe_params p;
};
void launch_debug_cascade(const DebugCascadeArgs& a, hipStream_t s);

}  // namespace swe
