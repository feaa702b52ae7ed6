// swe_step_kernel.hpp -- the tile-owned step kernel for one-call snapshot
// windows, templated on the tile shape (SX x SY cells, NT threads, a
// register budget of WPE waves per SIMD); each shape is instantiated in its
// own translation unit (swe_step_shapes.hip).
//
// With window == 1 every fly() call writes only to the particle's pre-move
// cell, i.e. the cell it occupies when the window starts.  k_step gives each
// workgroup one tile of kSX x kSY cells and the particles currently in it,
// so all of a window's writes land in LDS of the owning workgroup and reach
// HBM as plain coalesced stores -- no global atomics.  One launch per window:
//
//   1. relax of the previous window (the k_relax stencil) for the tile plus
//      a one-cell halo, recomputed redundantly from a three-cell snapshot
//      halo in LDS: master/snapshot of the tile written out, the halo kept in
//      LDS for the particles' five-point stencils;
//   2. one fly() call (wind.hpp:28-117) for every particle in the tile
//      against the LDS snapshot; pickups, cascade-event counts and tracks
//      summed in LDS (int64 / uint32) and stored once per cell;
//   3. routing: the tile's output region of the particle buffer has nine
//      sub-ranges, one per direction (-1, 0, +1)^2 of the tile a particle now
//      occupies, each with room for all the tile's particles; a particle is
//      appended to the sub-range of its direction.  The per-direction counts
//      are added to the destination tiles' input counts for the next window
//      (and to the sums of chunks of kChunk tiles), from which every
//      workgroup derives the start of its own output region.  A tile reads
//      its arrivals from the nine sub-ranges that point at it: no scan.
//
// A particle that jumps more than one tile (never seen in practice) goes to a
// "far" list that every workgroup scans for particles in its tile.
#pragma once

#include "swe_cascade.hpp"
#include "swe_fly.hpp"
#include "swe_kernels.h"
#include "swe_step.h"

namespace swe {

constexpr int kChunk = kStepChunk;   // tiles per input-count chunk sum
constexpr uint32_t kCrowded = 2048;  // tile load counted as crowded in the statistics

constexpr int32_t c_ox[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
constexpr int32_t c_oy[8] = {-1, 0, 1, -1, 1, -1, 0, 1};

// LDS-backed stencil and accumulator for one tile (kH1Y: row length of the
// one-cell halo, kCell: cells of the tile)
template <int kH1Y>
struct TileMap {
  const float* s_new;   // H1 heights
  const float* s_snew;  // H1 sediment
  int32_t c;            // H1 index of the particle's cell
  __device__ float h(int dx, int dy) const { return s_new[c + dx * kH1Y + dy]; }
  __device__ float s0() const { return s_snew[c]; }
};

template <int kCell>
struct TileAcc {
  int64_t* d;
  int64_t* ds;
  uint32_t* v;
  int64_t* t;
  int32_t li;  // tile-local index of the particle's cell
  __device__ void add_pickup(int64_t, int64_t q, bool both) {
    atomicAdd((unsigned long long*)&d[li], (unsigned long long)q);
    if (both) atomicAdd((unsigned long long*)&ds[li], (unsigned long long)q);
  }
  __device__ void add_event(int64_t) { atomicAdd(&v[li], 1u); }
  __device__ void add_track(int64_t, float sed, f3 sp) {
    atomicAdd((unsigned long long*)&t[li], (unsigned long long)fx_from_f32(sed));
    atomicAdd((unsigned long long*)&t[kCell + li], (unsigned long long)fx_from_f32(sp.x));
    atomicAdd((unsigned long long*)&t[2 * kCell + li], (unsigned long long)fx_from_f32(sp.y));
    atomicAdd((unsigned long long*)&t[3 * kCell + li], (unsigned long long)fx_from_f32(sp.z));
  }
};

// a copy of a uniform float in a vector register
__device__ __forceinline__ float in_vgpr(float x) {
  float y;
  asm volatile("v_mov_b32 %0, %1" : "=v"(y) : "s"(x));
  return y;
}

__device__ __forceinline__ int32_t wrapt(int32_t v, int32_t n) {
  // small-range wrap for tile coordinates (|v| < 2n)
  if (v < 0) v += n;
  if (v >= n) v -= n;
  return v;
}

// the register budget keeps WPE waves per SIMD (the default shape: five, and
// with the LDS of one workgroup under 32 KiB five workgroups share a CU)
// MODE 0: relax of the previous window (synchronous cascade) + one fly()
// call per particle; MODE 1: the relax only (the cycle's last window);
// MODE 2: the fly() calls only -- the previous window was relaxed before the
// launch by the colour-ordered cascade (one launch per colour class), so the
// tile reads the snapshot and masters as they are.
template <int SX, int SY, int NT, int WPE, bool PERIODIC, bool TRACK, int LAYERS, bool DEPOSIT, int MODE>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE))) void k_step(StepArgs a) {
  constexpr bool RELAX_ONLY = MODE == 1, EXT = MODE == 2;
  constexpr int kSX = SX, kSY = SY;
  constexpr int kH3X = kSX + 6, kH3Y = kSY + 6;  // snapshot halo 3
  constexpr int kH2X = kSX + 4, kH2Y = kSY + 4;  // events halo 2
  constexpr int kH1X = kSX + 2, kH1Y = kSY + 2;  // new values halo 1
  constexpr int kStepThreads = NT;
  constexpr int kCell = kSX * kSY;
  __shared__ float s_new[kH1X * kH1Y];
  __shared__ float s_snew[LAYERS == 2 ? kH1X * kH1Y : 1];
  // relax-phase and march-phase working sets share one buffer
  constexpr int kRelaxBytes =
      4 * kH3X * kH3Y * (LAYERS == 2 ? 2 : 1) + 4 * kH2X * kH2Y + 8 * kH1X * kH1Y + 2 * kH2X * kH2Y;
  constexpr int kMarchBytes = 8 * kCell * (LAYERS == 2 ? 2 : 1) + 4 * kCell + (TRACK ? 32 * kCell : 0);
  constexpr int kWorkBytes = kRelaxBytes > kMarchBytes ? kRelaxBytes : kMarchBytes;
  __shared__ __attribute__((aligned(16))) unsigned char s_work[kWorkBytes];
  __shared__ uint32_t s_slot[9];     // particles sent in direction d
  __shared__ int64_t s_rbase[9];     // arrivals: sub-range d of the source tile
  __shared__ uint32_t s_rpre[10];    // exclusive prefix of the sub-range lengths
  __shared__ int64_t s_lost[kStepThreads / 64];
  __shared__ uint64_t s_wstat[kStepThreads / 64];  // per wave: far moves << 32 | steps
  __shared__ int64_t s_base;
  __shared__ uint32_t s_ncas;        // cells whose cascade moves material

  // XCD-aware order (workgroup b runs on XCD b % 8).  tile_order 0: XCD k
  // takes a contiguous range of tiles, so neighbouring tiles share halo
  // lines in one L2.  tile_order C >= 1: runs of C consecutive tiles are
  // dealt round-robin over the XCDs (evens out crowded regions); the last
  // nb mod 8C workgroups keep their own index.  Both are bijections.
  // The first heavy_slots workgroups take the tiles listed as crowded by the
  // previous launch, the last light_slots ones the tiles listed as light
  // (at most kLightMax arrivals), both listed per home XCD (the XCD that
  // takes the tile in the usual order), so a listed tile keeps its L2;
  // empty slots return at once.  The rest follow the usual order and skip
  // a listed tile.  heavy_slots and default_slots (with light slots) are
  // multiples of 8, so slot b of each part runs on XCD b % 8.
  const uint32_t hs = RELAX_ONLY ? 0u : (uint32_t)a.heavy_slots;
  const uint32_t ls = RELAX_ONLY ? 0u : (uint32_t)a.light_slots;
  const uint32_t nb = (uint32_t)a.default_slots, b = blockIdx.x - (uint32_t)a.heavy_slots;
  const uint32_t xcd = b % 8;
  // a single map's terrain record does not depend on the tile: its load goes
  // out with the slot's list or stamp loads below, not after them
  const Terrain T1 = a.terr[0];
  int32_t u;  // global tile
  bool listed = false;
  if (blockIdx.x < (uint32_t)a.heavy_slots) {
    const uint32_t k = blockIdx.x / 8;  // heavy slot k of this XCD (blockIdx.x % 8 == xcd)
    if (hs == 0) return;
    // the list entry is read together with the list length (k < hs / 8, so
    // it lies inside the list whatever the length; unused when k >= length)
    const uint32_t cnt = a.hcnt_cur[xcd];
    const uint32_t cand = a.heavy_cur[xcd * (hs / 8) + k];
    if (k >= cnt) return;
    u = (int32_t)cand;
    listed = true;
  } else if (b >= nb) {
    const uint32_t bl = b - nb, k = bl / 8;  // k < ls / 8
    if (ls == 0) return;
    const uint32_t cnt = a.hcnt_cur[8 + xcd];
    const uint32_t cand = a.heavy_cur[kHeavySlotsMax + xcd * (ls / 8) + k];
    if (k >= cnt) return;
    u = (int32_t)cand;
    listed = true;
  } else if (a.tile_order <= 0) {
    const uint32_t q8 = nb / 8, r8 = nb % 8;
    u = (int32_t)(xcd * q8 + (xcd < r8 ? xcd : r8) + b / 8);
  } else {
    const uint32_t C = (uint32_t)a.tile_order, full = nb / (8 * C) * (8 * C), j = b / 8;
    u = b < full ? (int32_t)(((j / C) * 8 + xcd) * C + j % C) : (int32_t)b;
  }
  if (u >= a.ntiles) return;  // padding of the usual order to a multiple of 8
  // home XCD of tile t in the usual order: the XCD whose workgroups take it
  auto home_of = [=](uint32_t t) __attribute__((always_inline)) -> uint32_t {
    if (a.tile_order <= 0) {
      const uint32_t q8 = nb / 8, r8 = nb % 8;
      return t < r8 * (q8 + 1) ? t / (q8 + 1) : r8 + (t - r8 * (q8 + 1)) / q8;
    }
    const uint32_t C = (uint32_t)a.tile_order, full = nb / (8 * C) * (8 * C);
    return t < full ? (t / C) % 8 : t % 8;
  };
  // a listed tile's stamp is window + 1 (or window + 2 once its listed
  // workgroup has listed it again): either way the usual one skips it,
  // before it loads anything
  if ((hs | ls) > 0 && !listed && a.hstamp[u] > (uint32_t)a.window) return;
  // the terrain this tile belongs to; all map indices below are terrain-local
  Terrain T = T1;
  if (a.nterr != 1) T = a.terr[a.tile_terrain[u]];
  const int32_t nx = T.nx, ny = T.ny;
  const int32_t tiles_x = T.tiles_x, tiles_y = T.tiles_y;
  const int32_t t0 = T.tile0;
  const int32_t ul = u - t0;
  const int32_t tx = ul / tiles_y, ty = ul % tiles_y;
  const int32_t x0 = tx * kSX, y0 = ty * kSY;
  const float rcp_nx = T.rcp_nx, rcp_ny = T.rcp_ny;
  const int64_t c0 = T.cell0;
  const int64_t* zm_in = a.zm_in + c0;
  int64_t* zm_out = a.zm_out + c0;
  const float* snap_in = a.snap_in + c0;
  float* snap_out = a.snap_out + c0;
  const int64_t* sm_in = LAYERS == 2 ? a.sm_in + c0 : nullptr;
  int64_t* sm_out = LAYERS == 2 ? a.sm_out + c0 : nullptr;
  const float* ssnap_in = LAYERS == 2 ? a.ssnap_in + c0 : nullptr;
  float* ssnap_out = LAYERS == 2 ? a.ssnap_out + c0 : nullptr;
  const uint32_t* v_in = a.v_in + c0;
  uint32_t* v_out = a.v_out + c0;
  int64_t* trk = TRACK ? a.trk + c0 : nullptr;

  constexpr int kCellPer = (kCell + kStepThreads - 1) / kStepThreads;

  // this window's output region starts at 9 x (input counts of tiles 0 ..
  // u-1): the chunk sums of the earlier chunks plus the earlier tiles of this
  // chunk.  The loads are issued here and summed after the relax phase (no
  // wait on them before the map loads go out).
  static_assert(kChunk <= kStepThreads, "one tile of the current chunk per thread");
  uint32_t part_c = 0, part_n = 0;
  {
    const int32_t cu = u / kChunk;
    if (cu <= kStepThreads) {
      if ((int32_t)threadIdx.x < cu) part_c = a.csum_cur[threadIdx.x];
    } else {
      for (int32_t c = threadIdx.x; c < cu; c += kStepThreads) part_c += a.csum_cur[c];
    }
    const int32_t t = cu * kChunk + (int32_t)threadIdx.x;
    if (t < u) part_n = a.nin_cur[t];
  }
  const uint32_t nin_u = RELAX_ONLY ? 0u : a.nin_cur[u];

  PRec rec0{};       // this thread's first arrival (loaded in the relax phase)
  uint32_t sr0 = 0;  // its sub-range
  if constexpr (EXT) {
    // ---------------- 1'. the relaxed map of the previous window ----------------
    // one-halo snapshot into LDS for the particles' stencils, the tile's
    // masters into the pickup accumulators; all loads issued first
    constexpr int kH1Per = (kH1X * kH1Y + kStepThreads - 1) / kStepThreads;
    float hv[kH1Per], hs1[LAYERS == 2 ? kH1Per : 1];
    int64_t zt[kCellPer], st[LAYERS == 2 ? kCellPer : 1];
#pragma unroll
    for (int i = 0; i < kH1Per; i++) {
      const int e = threadIdx.x + i * kStepThreads;
      int32_t gx = x0 - 1 + e / kH1Y, gy = y0 - 1 + e % kH1Y;
      bool ok = e < kH1X * kH1Y;
      if (PERIODIC) {
        gx = wrap(gx, nx, rcp_nx);
        gy = wrap(gy, ny, rcp_ny);
      } else {
        ok = ok && gx >= 0 && gy >= 0 && gx < nx && gy < ny;
      }
      const int64_t g = (int64_t)gx * ny + gy;
      hv[i] = ok ? snap_in[g] : 0.0f;
      if (LAYERS == 2) hs1[i] = ok ? ssnap_in[g] : 0.0f;
    }
#pragma unroll
    for (int i = 0; i < kCellPer; i++) {
      const int e = threadIdx.x + i * kStepThreads;
      const int32_t gx = x0 + e / kSY, gy = y0 + e % kSY;
      const bool ok = e < kCell && gx < nx && gy < ny;
      const int64_t g = (int64_t)gx * ny + gy;
      zt[i] = ok ? zm_in[g] : 0;
      if (LAYERS == 2) st[i] = ok ? sm_in[g] : 0;
    }
    if (threadIdx.x < 9) {
      const int32_t d = threadIdx.x, dx = d / 3 - 1, dy = d % 3 - 1;
      int32_t sx = tx - dx, sy = ty - dy;
      bool ok = true;
      if (PERIODIC) {
        sx = wrapt(sx, tiles_x);
        sy = wrapt(sy, tiles_y);
      } else {
        ok = sx >= 0 && sy >= 0 && sx < tiles_x && sy < tiles_y;
      }
      int64_t rb = 0;
      uint32_t rc = 0;
      if (ok) {
        const int64_t s9 = (int64_t)(t0 + sx * tiles_y + sy) * 9 + d;
        rb = a.sub_base_prev[s9];
        rc = a.sub_cnt_prev[s9];
      }
      s_rbase[d] = rb;
      s_slot[d] = rc;
    }
    int64_t* acc_z = (int64_t*)s_work;
#pragma unroll
    for (int i = 0; i < kH1Per; i++) {
      const int e = threadIdx.x + i * kStepThreads;
      if (e < kH1X * kH1Y) {
        s_new[e] = hv[i];
        if (LAYERS == 2) s_snew[e] = hs1[i];
      }
    }
#pragma unroll
    for (int i = 0; i < kCellPer; i++) {
      const int e = threadIdx.x + i * kStepThreads;
      if (e < kCell) {
        acc_z[e] = zt[i];
        if (LAYERS == 2) acc_z[kCell + e] = st[i];
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t r = 0;
      for (int d = 0; d < 9; d++) {
        s_rpre[d] = r;
        r += s_slot[d];
      }
      s_rpre[9] = r;
    }
    __syncthreads();
    if (threadIdx.x < s_rpre[9]) {
      while (sr0 < 8 && threadIdx.x >= s_rpre[sr0 + 1]) sr0++;
      rec0 = a.in[s_rbase[sr0] + (int64_t)(threadIdx.x - s_rpre[sr0])];
    }
  } else {
  // ---------------- 1. relax of the previous window ----------------
  {
    float* s_old = (float*)s_work;
    float* s_sold = s_old + kH3X * kH3Y;  // valid when LAYERS == 2
    uint32_t* s_vold = (uint32_t*)(s_old + kH3X * kH3Y * (LAYERS == 2 ? 2 : 1));
    int64_t* s_casc = (int64_t*)(s_vold + kH2X * kH2Y);
    // master (with the previous window's pickups) of the tile and its halo,
    // loaded before the cascades are computed
    constexpr int kH1Per = (kH1X * kH1Y + kStepThreads - 1) / kStepThreads;
    int64_t zpre[kH1Per], spre[LAYERS == 2 ? kH1Per : 1];
#pragma unroll
    for (int i = 0; i < kH1Per; i++) {
      const int e = threadIdx.x + i * kStepThreads;
      int32_t gx = x0 - 1 + e / kH1Y, gy = y0 - 1 + e % kH1Y;
      bool ok = e < kH1X * kH1Y;
      if (PERIODIC) {
        gx = wrap(gx, nx, rcp_nx);
        gy = wrap(gy, ny, rcp_ny);
      } else {
        ok = ok && gx >= 0 && gy >= 0 && gx < nx && gy < ny;
      }
      const int64_t g = (int64_t)gx * ny + gy;
      zpre[i] = ok ? zm_in[g] : 0;
      if (LAYERS == 2) spre[i] = ok ? sm_in[g] : 0;
    }
    // snapshot (3-halo) and events (2-halo): every load is issued before the
    // first LDS store, so the fill costs one memory round trip
    constexpr int kH3Per = (kH3X * kH3Y + kStepThreads - 1) / kStepThreads;
    constexpr int kH2Per = (kH2X * kH2Y + kStepThreads - 1) / kStepThreads;
    float hv[kH3Per], hs3[LAYERS == 2 ? kH3Per : 1];
    uint32_t vv[kH2Per];
#pragma unroll
    for (int i = 0; i < kH3Per; i++) {
      const int e = threadIdx.x + i * kStepThreads;
      int32_t gx = x0 - 3 + e / kH3Y, gy = y0 - 3 + e % kH3Y;
      bool ok = e < kH3X * kH3Y;
      if (PERIODIC) {
        gx = wrap(gx, nx, rcp_nx);
        gy = wrap(gy, ny, rcp_ny);
      } else {
        ok = ok && gx >= 0 && gy >= 0 && gx < nx && gy < ny;
      }
      const int64_t g = (int64_t)gx * ny + gy;
      hv[i] = ok ? snap_in[g] : 0.0f;
      if (LAYERS == 2) hs3[i] = ok ? ssnap_in[g] : 0.0f;
    }
#pragma unroll
    for (int i = 0; i < kH2Per; i++) {
      const int e = threadIdx.x + i * kStepThreads;
      int32_t gx = x0 - 2 + e / kH2Y, gy = y0 - 2 + e % kH2Y;
      bool ok = e < kH2X * kH2Y;
      if (PERIODIC) {
        gx = wrap(gx, nx, rcp_nx);
        gy = wrap(gy, ny, rcp_ny);
      } else {
        ok = ok && gx >= 0 && gy >= 0 && gx < nx && gy < ny;
      }
      vv[i] = ok ? v_in[(int64_t)gx * ny + gy] : 0u;
    }
    // arrivals of this window: the source tile in direction -d of this one
    // wrote the particles that moved by d into its sub-range d (loaded after
    // the map loads are out, so that their wait overlaps those)
    if (threadIdx.x < 9) {
      const int32_t d = threadIdx.x, dx = d / 3 - 1, dy = d % 3 - 1;
      int32_t sx = tx - dx, sy = ty - dy;
      bool ok = !RELAX_ONLY;
      if (PERIODIC) {
        sx = wrapt(sx, tiles_x);
        sy = wrapt(sy, tiles_y);
      } else {
        ok = ok && sx >= 0 && sy >= 0 && sx < tiles_x && sy < tiles_y;
      }
      int64_t rb = 0;
      uint32_t rc = 0;
      if (ok) {
        const int64_t s9 = (int64_t)(t0 + sx * tiles_y + sy) * 9 + d;
        rb = a.sub_base_prev[s9];
        rc = a.sub_cnt_prev[s9];
      }
      s_rbase[d] = rb;
      s_slot[d] = rc;  // lengths, turned into s_rpre below
    }
#pragma unroll
    for (int i = 0; i < kH3Per; i++) {
      const int e = threadIdx.x + i * kStepThreads;
      if (e < kH3X * kH3Y) {
        s_old[e] = hv[i];
        if (LAYERS == 2) s_sold[e] = hs3[i];
      }
    }
#pragma unroll
    for (int i = 0; i < kH2Per; i++) {
      const int e = threadIdx.x + i * kStepThreads;
      if (e < kH2X * kH2Y) s_vold[e] = vv[i];
    }
    uint16_t* s_cas = (uint16_t*)(s_casc + kH1X * kH1Y);  // H2 cells with a transfer
    for (int e = threadIdx.x; e < kH1X * kH1Y; e += kStepThreads) s_casc[e] = 0;
    if (threadIdx.x == 0) s_ncas = 0;
    __syncthreads();
    // start of each arrival sub-range in this tile's arrival order (read
    // after the next barrier)
    if (threadIdx.x == 0) {
      uint32_t r = 0;
      for (int d = 0; d < 9; d++) {
        s_rpre[d] = r;
        r += s_slot[d];
      }
      s_rpre[9] = r;
    }

    // visited cells within two cells of the tile whose cascade moves
    // material: one neighbour beyond the threshold at the window's start
    // height.  The walk changes the centre only after a transfer, so a cell
    // with none such moves nothing and is skipped.  Listed in LDS so that
    // the cascades below run on full waves.
    for (int e0 = 0; e0 < kH2X * kH2Y; e0 += kStepThreads) {
      const int e = e0 + (int)threadIdx.x;
      bool act = false;
      if (e < kH2X * kH2Y && s_vold[e] != 0) {
        const int32_t lx = e / kH2Y, ly = e % kH2Y;
        const int32_t gx = x0 - 2 + lx, gy = y0 - 2 + ly;
        const int32_t ic = (lx + 1) * kH3Y + (ly + 1);
        const float hc = s_old[ic];
        const float dd = __builtin_sqrtf(2.0f);
#pragma unroll
        for (int k = 0; k < 8; k++) {
          bool ok = true;
          if (!PERIODIC) {
            const int32_t qx = gx + c_ox[k], qy = gy + c_oy[k];
            ok = qx >= 0 && qy >= 0 && qx < nx && qy < ny;
          }
          const float diff = hc - s_old[ic + c_ox[k] * kH3Y + c_oy[k]];
          const float d = (c_ox[k] != 0 && c_oy[k] != 0) ? dd : 1.0f;
          const float excess = __builtin_fabsf(diff) - (d * a.maxdiff) * 1.0f;
          act = act || (ok && diff != 0.0f && !(excess <= 0.0f));
        }
      }
      const uint64_t m = __ballot(act);
      const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
      uint32_t base = 0;
      if ((threadIdx.x & 63) == 0 && m) base = atomicAdd(&s_ncas, (uint32_t)__popcll(m));
      base = __shfl(base, 0, 64);
      if (act) s_cas[base + below] = (uint16_t)e;
    }
    __syncthreads();

    // their cascades
    const uint32_t ncas = s_ncas;
    for (uint32_t i = threadIdx.x; i < ncas; i += kStepThreads) {
      const int32_t e = s_cas[i];
      const uint32_t v = s_vold[e];
      const int32_t lx = e / kH2Y, ly = e % kH2Y;  // H2 coords; H3 = +1, H1 = -1
      const int32_t gx = x0 - 2 + lx, gy = y0 - 2 + ly;
      float hn[8];
      uint32_t valid = 0;
      const int32_t ic = (lx + 1) * kH3Y + (ly + 1);
#pragma unroll
      for (int k = 0; k < 8; k++) {
        hn[k] = s_old[ic + c_ox[k] * kH3Y + c_oy[k]];
        bool ok = true;
        if (!PERIODIC) {
          const int32_t qx = gx + c_ox[k], qy = gy + c_oy[k];
          ok = qx >= 0 && qy >= 0 && qx < nx && qy < ny;
        }
        valid |= (ok ? 1u : 0u) << k;
      }
      const auto sed_of = [&](int32_t k) { return s_sold[ic + off_x(k) * kH3Y + off_y(k)]; };
      int64_t centre = 0;
      cascade_cell_dev<LAYERS>(hn, sed_of, s_old[ic], LAYERS == 2 ? s_sold[ic] : 0.0f, valid,
                               settle_factor_dev(a.settling, v), a.maxdiff, [&](int32_t k, int64_t amt) {
                                 centre -= amt;
                                 const int32_t hx = lx - 1 + off_x(k), hy = ly - 1 + off_y(k);
                                 if (amt != 0 && hx >= 0 && hy >= 0 && hx < kH1X && hy < kH1Y)
                                   atomicAdd((unsigned long long*)&s_casc[hx * kH1Y + hy], (unsigned long long)amt);
                               });
      const int32_t hx = lx - 1, hy = ly - 1;
      if (centre != 0 && hx >= 0 && hy >= 0 && hx < kH1X && hy < kH1Y)
        atomicAdd((unsigned long long*)&s_casc[hx * kH1Y + hy], (unsigned long long)centre);
    }
    // each thread's first arrival record, loaded now so that it arrives
    // while the other threads finish their cascades and the relax ends
    if (!RELAX_ONLY && threadIdx.x < s_rpre[9]) {
      while (sr0 < 8 && threadIdx.x >= s_rpre[sr0 + 1]) sr0++;
      rec0 = a.in[s_rbase[sr0] + (int64_t)(threadIdx.x - s_rpre[sr0])];
    }
    __syncthreads();

    // master += cascade; next snapshot for the tile and its halo.  The
    // tile's own masters seed the particle phase's pickup accumulators
    // (acc_d / acc_ds below: the front of s_work, whose relax-phase arrays
    // are dead once the cascades are summed, and clear of s_casc)
    static_assert(8 * kCell * (LAYERS == 2 ? 2 : 1) <= 4 * kH3X * kH3Y * (LAYERS == 2 ? 2 : 1) + 4 * kH2X * kH2Y,
                  "the pickup accumulators must not overlap s_casc");
    int64_t* acc_z = (int64_t*)s_work;
#pragma unroll
    for (int i = 0; i < kH1Per; i++) {
      const int e = threadIdx.x + i * kStepThreads;
      if (e >= kH1X * kH1Y) continue;
      const int32_t lx = e / kH1Y, ly = e % kH1Y;
      int32_t gx = x0 - 1 + lx, gy = y0 - 1 + ly;
      bool ok = true;
      if (PERIODIC) {
        gx = wrap(gx, nx, rcp_nx);
        gy = wrap(gy, ny, rcp_ny);
      } else {
        ok = gx >= 0 && gy >= 0 && gx < nx && gy < ny;
      }
      if (!ok) {
        s_new[e] = 0.0f;
        if (LAYERS == 2) s_snew[e] = 0.0f;
        continue;
      }
      const int64_t g = (int64_t)gx * ny + gy;
      const int64_t cas = s_casc[e];
      const int64_t z = zpre[i] + cas;
      const float zf = f32_from_fx(z);
      s_new[e] = zf;
      const bool mine = lx >= 1 && ly >= 1 && lx <= kSX && ly <= kSY && x0 - 1 + lx < nx && y0 - 1 + ly < ny;
      const int32_t li = (lx - 1) * kSY + (ly - 1);
      if (mine) {
        acc_z[li] = z;
        snap_out[g] = zf;
      }
      if (LAYERS == 2) {
        int64_t s = spre[i] + cas;
        if (s < 0) s = 0;
        const float sf = f32_from_fx(s);
        s_snew[e] = sf;
        if (mine) {
          acc_z[kCell + li] = s;
          ssnap_out[g] = sf;
        }
      }
    }
    __syncthreads();
  }
  }  // relax (MODE 0, 1)
  {
    int64_t part = (int64_t)part_c + (int64_t)part_n;
    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
    if ((threadIdx.x & 63) == 0) s_lost[threadIdx.x >> 6] = part;
    __syncthreads();
    if (threadIdx.x == 0) {
      int64_t b0 = 0;
      for (int i = 0; i < kStepThreads / 64; i++) b0 += s_lost[i];
      s_base = 9 * b0;
    }
  }

  // ---------------- 2. particles of the tile ----------------
  int64_t* acc_d = (int64_t*)s_work;
  int64_t* acc_ds = acc_d + kCell;  // LAYERS == 2
  uint32_t* acc_v = (uint32_t*)(acc_d + kCell * (LAYERS == 2 ? 2 : 1));
  int64_t* acc_t = (int64_t*)(acc_v + kCell);  // TRACK (kCell*4), 8-byte aligned: kCell*4 bytes of acc_v
  for (int e = threadIdx.x; e < kCell; e += kStepThreads) {
    acc_v[e] = 0;  // acc_d / acc_ds hold the masters, the pickups add to them
    if (TRACK) {
      acc_t[e] = 0;
      acc_t[kCell + e] = 0;
      acc_t[2 * kCell + e] = 0;
      acc_t[3 * kCell + e] = 0;
    }
  }
  __syncthreads();  // s_rpre is read from s_slot before s_slot restarts
  if (threadIdx.x < 9) s_slot[threadIdx.x] = 0;
  int64_t lost_here = 0;
  uint32_t steps_here = 0, far_here = 0;
  __syncthreads();

  if (!RELAX_ONLY) {
    FlyConsts k;
    k.nx = nx;
    k.ny = ny;
    k.rcp_nx = rcp_nx;
    k.rcp_ny = rcp_ny;
    k.mapscale = a.mapscale;
    k.max_age = a.max_age;
    set_boundarylayer(k, a.boundarylayer);
    k.suspension = a.suspension;
    k.gravity = a.gravity;
    k.abrasion = a.abrasion;
    k.ps = mk(T.ps[0], T.ps[1], T.ps[2]);
    k.nps = mk(T.nps[0], T.nps[1], T.nps[2]);
    // uniform constants of fly() held in vector registers (one each) rather
    // than scalar ones, which this kernel has too few of
    k.mapscale = in_vgpr(k.mapscale);
    k.boundarylayer = in_vgpr(k.boundarylayer);
    k.gravity = in_vgpr(k.gravity);
    if (!EXT) {  // in the step-only mode these stay scalar: pinning them too spills vector registers there
      k.rcp_bl = in_vgpr(k.rcp_bl);
      k.suspension = in_vgpr(k.suspension);
      k.abrasion = in_vgpr(k.abrasion);
    }
    k.ps = mk(in_vgpr(k.ps.x), in_vgpr(k.ps.y), in_vgpr(k.ps.z));
    k.nps = mk(in_vgpr(k.nps.x), in_vgpr(k.nps.y), in_vgpr(k.nps.z));
    const uint64_t ck = cycle_key(T.seed, *a.cycle);
    const int64_t p0 = T.part0;
    const int64_t out_base = s_base;  // 9 x the input counts of tiles 0 .. u-1
    const uint32_t far_n = *a.far_cnt_in;

    // one input record
    using Rec = PRec;
    auto load = [](const PRec* in, int64_t j) __attribute__((always_inline)) { return in[j]; };
    // a particle more than one tile away goes to the far list; the tile of
    // cell (fx, fy) counts it among its next window's arrivals
    auto send_far = [&](f4a8 o0, f4a8 o1, int64_t ledger, int32_t fx, int32_t fy) __attribute__((always_inline)) {
      const int32_t dest = t0 + (fx / kSX) * tiles_y + fy / kSY;
      atomicAdd(&a.nin_next[dest], 1u);
      atomicAdd(&a.csum_next[dest / kChunk], 1u);
      const uint32_t fslot = atomicAdd(a.far_cnt_out, 1u);
      a.far_out[fslot] = PRec{o0, o1, ledger};
    };
    // step one record and write it out
    auto process = [&](const Rec& r) __attribute__((always_inline)) {
      Particle w;
      const f4a8 q0 = r.q0;
      const f4a8 q1 = r.q1;
      w.px = q0.x; w.py = q0.y; w.pz = q0.z; w.sed = q0.w;
      w.sx = q1.x; w.sy = q1.y; w.sz = q1.z;
      w.age = a.window;
      w.ledger = r.ledger;
      w.lastc = 0;
      const uint32_t pid = f2u(q1.w);
      const uint64_t pk = particle_key(ck, (uint64_t)((int64_t)pid - p0));
      int32_t cx, cy;
      const bool ok = cell_of<PERIODIC>(w.px, w.pz, k, cx, cy);
      int64_t dead_cell = -1;
      bool dead = !ok;
      if (ok && (cx < x0 || cy < y0 || cx >= x0 + kSX || cy >= y0 + kSY)) {
        // routing always places a particle in the tile of its cell; keep
        // LDS indices in range whatever happens by forwarding a stray one
        send_far(q0, q1, w.ledger, cx, cy);
        return;
      }
      if (ok) {
        const int32_t lx = cx - x0, ly = cy - y0;
        TileMap<kH1Y> m{s_new, s_snew, (lx + 1) * kH1Y + (ly + 1)};
        TileAcc<kCell> acc{acc_d, acc_ds, acc_v, acc_t, lx * kSY + ly};
        if (fly<PERIODIC, TRACK, LAYERS>(w, pk, cx, cy, k, m, acc) == kDied) {
          dead = true;
          dead_cell = (int64_t)cx * ny + cy;
        }
      }
      if (dead) {
        // died in this call at its cell, or found off the map: the cell it left
        const int64_t c = dead_cell >= 0 ? dead_cell : (int64_t)f2u(q0.x);
        if (DEPOSIT) {
          const int32_t lx = (int32_t)(c / ny) - x0, ly = (int32_t)(c % ny) - y0;
          if (w.ledger != 0 && lx >= 0 && ly >= 0 && lx < kSX && ly < kSY) {
            const int32_t li = lx * kSY + ly;
            atomicAdd((unsigned long long*)&acc_d[li], (unsigned long long)w.ledger);
            if (LAYERS == 2) atomicAdd((unsigned long long*)&acc_ds[li], (unsigned long long)w.ledger);
          }
        } else {
          lost_here += w.ledger;
        }
        return;
      }
      steps_here++;
      // route by the cell the particle now occupies; one that leaves a
      // bounded map (or whose position stops being finite) stays with the
      // tile of its last cell and dies there at its next call
      int32_t ncx = 0, ncy = 0;
      int32_t code = 4;
      bool off = true;
      if (cell_of<PERIODIC>(w.px, w.pz, k, ncx, ncy)) {
        off = false;
        int32_t dtx = ncx / kSX - tx, dty = ncy / kSY - ty;
        if (PERIODIC) {
          if (tiles_x > 2) {
            if (dtx == tiles_x - 1) dtx = -1;
            if (dtx == 1 - tiles_x) dtx = 1;
          }
          if (tiles_y > 2) {
            if (dty == tiles_y - 1) dty = -1;
            if (dty == 1 - tiles_y) dty = 1;
          }
        }
        code = (dtx < -1 || dtx > 1 || dty < -1 || dty > 1) ? 255 : (dtx + 1) * 3 + (dty + 1);
      }
      const f4a8 o0 = off ? f4a8{u2f((uint32_t)w.lastc), w.py, __builtin_nanf(""), w.sed}
                          : f4a8{w.px, w.py, w.pz, w.sed};
      const f4a8 o1 = {w.sx, w.sy, w.sz, u2f(pid)};
      if (code == 255) {
        far_here++;
        send_far(o0, o1, w.ledger, ncx, ncy);
        return;
      }
      // sub-range `code` of this tile's output region holds up to nin_u records
      const int64_t o = out_base + (int64_t)code * nin_u + atomicAdd(&s_slot[code], 1u);
      a.out[o] = PRec{o0, o1, w.ledger};
    };

    // the arrivals, e = 0 .. n_in - 1 over the nine sub-ranges in order; each
    // thread's e increase, so its sub-range cursor only moves forward.  The
    // next record is loaded while the current one is stepped.
    const uint32_t n_in = s_rpre[9];
    uint32_t sr = sr0;
    auto index_of = [&](uint32_t e) __attribute__((always_inline)) {
      while (sr < 8 && e >= s_rpre[sr + 1]) sr++;
      return s_rbase[sr] + (int64_t)(e - s_rpre[sr]);
    };
    {
      Rec cur = rec0;
      for (uint32_t e = threadIdx.x; e < n_in; e += kStepThreads) {
        Rec nxt{};
        if (e + kStepThreads < n_in) nxt = load(a.in, index_of(e + kStepThreads));
        process(cur);
        cur = nxt;
      }
    }
    for (int64_t j = threadIdx.x; j < (int64_t)far_n; j += kStepThreads) {
      const int64_t fp = (int64_t)f2u(a.far_in[j].q1.w);
      if (fp < p0 || fp >= p0 + T.npart) continue;  // another terrain's particle
      int32_t cx, cy;
      const f4a8 q0 = a.far_in[j].q0;
      int32_t ctx, cty;
      if (cell_of<PERIODIC>(q0.x, q0.z, k, cx, cy)) {
        ctx = cx / kSX;
        cty = cy / kSY;
      } else {
        const int32_t lc = (int32_t)f2u(q0.x);
        ctx = (lc / ny) / kSX;
        cty = (lc % ny) / kSY;
      }
      if (ctx == tx && cty == ty) process(load(a.far_in, j));
    }
  }
  __syncthreads();

  // ---------------- 3. write the tile's window results ----------------
#pragma unroll
  for (int i = 0; i < kCellPer; i++) {
    const int e = threadIdx.x + i * kStepThreads;
    const int32_t gx = x0 + e / kSY, gy = y0 + e % kSY;
    if (e >= kCell || gx >= nx || gy >= ny) continue;
    const int64_t g = (int64_t)gx * ny + gy;
    // master with this window's pickups; the clamp of the sediment layer at
    // zero follows the next window's cascades
    zm_out[g] = acc_d[e];
    v_out[g] = acc_v[e];
    if (LAYERS == 2) sm_out[g] = acc_ds[e];
    if (EXT) {  // the snapshot the colour-ordered relax starts from
      snap_out[g] = f32_from_fx(acc_d[e]);
      if (LAYERS == 2) ssnap_out[g] = f32_from_fx(acc_ds[e]);
    }
  }
  // flow tracks of the cells this window's particles visited, read here
  // rather than prefetched (16 fewer vector registers through the particle
  // loop lets five workgroups share a CU): the loads are issued now and the
  // sums stored at the very end, after the rest of the write-out
  int64_t tv[TRACK && !RELAX_ONLY ? kCellPer : 1][4], old[TRACK && !RELAX_ONLY ? kCellPer : 1][4];
  int64_t gi[TRACK && !RELAX_ONLY ? kCellPer : 1];
  if (TRACK && !RELAX_ONLY) {
#pragma unroll
    for (int i = 0; i < kCellPer; i++) {
      const int e = threadIdx.x + i * kStepThreads;
      const int32_t gx = x0 + e / kSY, gy = y0 + e % kSY;
      const bool in = e < kCell && gx < nx && gy < ny;
      gi[i] = (int64_t)gx * ny + gy;
#pragma unroll
      for (int f = 0; f < 4; f++) tv[i][f] = in ? acc_t[f * kCell + (in ? e : 0)] : 0;
    }
#pragma unroll
    for (int i = 0; i < kCellPer; i++)
#pragma unroll
      for (int f = 0; f < 4; f++) old[i][f] = tv[i][f] != 0 ? trk[f * a.ncell + gi[i]] : 0;
  }
  // s_rbase is free again (the arrivals are consumed): it carries the chunk
  // of each direction's destination tile to the aggregation below
  int32_t* s_dchunk = (int32_t*)s_rbase;
  if (threadIdx.x < 9) {
    // particles sent towards direction d: sub-range d of this tile's output,
    // and the next window's arrivals of the tile in direction d
    const int32_t d = threadIdx.x;
    const uint32_t n = s_slot[d];
    a.sub_base_cur[(int64_t)u * 9 + d] = s_base + (int64_t)d * nin_u;
    a.sub_cnt_cur[(int64_t)u * 9 + d] = n;
    int32_t chunk = -1;
    if (n) {
      const int32_t dx = d / 3 - 1, dy = d % 3 - 1;
      int32_t dtx = tx + dx, dty = ty + dy;
      if (PERIODIC) {
        dtx = wrapt(dtx, tiles_x);
        dty = wrapt(dty, tiles_y);
      }
      const int32_t dest = t0 + dtx * tiles_y + dty;
      atomicAdd(&a.nin_next[dest], n);
      chunk = dest / kChunk;
    }
    s_dchunk[d] = chunk;
  }
  if (threadIdx.x == 0) {
    a.nin_clear[u] = 0;  // the window after next accumulates here
    if (u % kChunk == 0) a.csum_clear[u / kChunk] = 0;
    if (u == 0) {
      *a.far_cnt_clear = 0;
      if ((hs | ls) > 0)
        for (int i = 0; i < 16; i++) a.hcnt_clear[i] = 0;
    }
    // crowded (or light) now: listed for the next launch (the count moves
    // little from one window to the next)
    if (hs > 0 && nin_u > kHeavyMin) {
      const uint32_t home = home_of((uint32_t)u);
      const uint32_t slot = atomicAdd(&a.hcnt_next[home], 1u);
      if (slot < hs / 8) {
        a.heavy_next[home * (hs / 8) + slot] = (uint32_t)u;
        a.hstamp[u] = (uint32_t)a.window + 2u;
      }
    } else if (ls > 0 && nin_u <= kLightMax) {
      const uint32_t home = home_of((uint32_t)u);
      const uint32_t slot = atomicAdd(&a.hcnt_next[8 + home], 1u);
      if (slot < ls / 8) {
        a.heavy_next[kHeavySlotsMax + home * (ls / 8) + slot] = (uint32_t)u;
        a.hstamp[u] = (uint32_t)a.window + 2u;
      }
    }
  }
  if (!DEPOSIT) {
    for (int off = 32; off > 0; off >>= 1) lost_here += __shfl_xor(lost_here, off, 64);
    if ((threadIdx.x & 63) == 0) s_lost[threadIdx.x >> 6] = lost_here;
  }
  if (!RELAX_ONLY) {
    for (int off = 32; off > 0; off >>= 1) {
      steps_here += __shfl_xor(steps_here, off, 64);
      far_here += __shfl_xor(far_here, off, 64);
    }
    if ((threadIdx.x & 63) == 0) s_wstat[threadIdx.x >> 6] = ((uint64_t)far_here << 32) | steps_here;
  }
  __syncthreads();
  if (!RELAX_ONLY && threadIdx.x < 9) {
    // chunk sums of the next window's arrivals: one atomic per distinct chunk
    // (the nine destinations mostly share one or two), by the thread of the
    // chunk's first direction
    const int d = threadIdx.x;
    const int32_t ch = s_dchunk[d];
    bool first = ch >= 0;
#pragma unroll
    for (int e = 0; e < 9; e++) first = first && !(e < d && s_dchunk[e] == ch);
    if (first) {
      uint32_t n = 0;
#pragma unroll
      for (int e = 0; e < 9; e++) n += e >= d && s_dchunk[e] == ch ? s_slot[e] : 0u;
      atomicAdd(&a.csum_next[ch], n);
    }
  }
  if (!RELAX_ONLY && threadIdx.x == 0) {
    // statistics: one atomic per workgroup and kind
    uint64_t steps = 0, far = 0;
    for (int i = 0; i < kStepThreads / 64; i++) {
      steps += (uint32_t)s_wstat[i];
      far += s_wstat[i] >> 32;
    }
    const uint32_t slot = (uint32_t)u & (kStatSlots - 1);
    if (steps) atomicAdd((unsigned long long*)&a.stat[kStatSteps * kStatSlots + slot], (unsigned long long)steps);
    if (far) atomicAdd((unsigned long long*)&a.stat[kStatFar * kStatSlots + slot], (unsigned long long)far);
    if (nin_u) atomicMax((unsigned long long*)&a.stat[kStatMaxTile * kStatSlots + slot], (unsigned long long)nin_u);
    if (nin_u > kCrowded) atomicAdd((unsigned long long*)&a.stat[kStatCrowded * kStatSlots + slot], 1ull);
    if (listed) {
      const bool light_slot = blockIdx.x >= (uint32_t)(a.heavy_slots + a.default_slots);
      atomicAdd((unsigned long long*)&a.stat[(light_slot ? kStatLight : kStatHeavy) * kStatSlots + slot], 1ull);
    }
  }
  __syncthreads();
  if (!DEPOSIT && threadIdx.x == 0) {
    int64_t l = 0;
    for (int i = 0; i < kStepThreads / 64; i++) l += s_lost[i];
    if (l != 0)
      atomicAdd((unsigned long long*)&a.lost[(int64_t)a.tile_terrain[u] * kLostSlots + (u & (kLostSlots - 1))],
                (unsigned long long)l);
  }
  if (TRACK && !RELAX_ONLY) {
#pragma unroll
    for (int i = 0; i < kCellPer; i++)
#pragma unroll
      for (int f = 0; f < 4; f++)
        if (tv[i][f] != 0) trk[f * a.ncell + gi[i]] = old[i][f] + tv[i][f];
  }
}

// ---------------------------------------------------------------------------
// Launchers of one shape (instantiated per shape in swe_step_shapes.hip)
// ---------------------------------------------------------------------------
template <int SX, int SY, int NT, int WPE>
struct StepLauncher {
  template <bool P, bool T, int L, bool D>
  static void step4(const StepArgs& a, int mode, hipStream_t s) {
    const unsigned g = (unsigned)(a.heavy_slots + a.default_slots + a.light_slots);
    if (mode == 1)
      hipLaunchKernelGGL((k_step<SX, SY, NT, WPE, P, T, L, D, 1>), dim3(g), dim3(NT), 0, s, a);
    else if (mode == 2)
      hipLaunchKernelGGL((k_step<SX, SY, NT, WPE, P, T, L, D, 2>), dim3(g), dim3(NT), 0, s, a);
    else
      hipLaunchKernelGGL((k_step<SX, SY, NT, WPE, P, T, L, D, 0>), dim3(g), dim3(NT), 0, s, a);
  }
  template <bool P, bool T, int L>
  static void step3(const StepArgs& a, int ro, bool dep, hipStream_t s) {
    if (dep) step4<P, T, L, true>(a, ro, s);
    else step4<P, T, L, false>(a, ro, s);
  }
  template <bool P, bool T>
  static void step2(const StepArgs& a, int ro, bool dep, int layers, hipStream_t s) {
    if (layers == 2) step3<P, T, 2>(a, ro, dep, s);
    else step3<P, T, 1>(a, ro, dep, s);
  }
  template <bool P>
  static void step1(const StepArgs& a, int ro, bool dep, bool trk, int layers, hipStream_t s) {
    if (trk) step2<P, true>(a, ro, dep, layers, s);
    else step2<P, false>(a, ro, dep, layers, s);
  }
  static void launch(const StepArgs& a, bool periodic, bool track, int layers, bool deposit, int mode,
                     hipStream_t s) {
    if (periodic) step1<true>(a, mode, deposit, track, layers, s);
    else step1<false>(a, mode, deposit, track, layers, s);
  }
};

}  // namespace swe
