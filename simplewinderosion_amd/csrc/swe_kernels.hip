// swe_kernels.hip -- gfx950 kernels of the erosion hot path.
//
//   k_march   one particle per lane: up to R Wind::fly() calls
//             (wind.hpp:28-117) against the window's float32 snapshot,
//             int64 fixed-point pickup deltas and per-cell cascade-event
//             counts by device-scope atomics.
//   k_relax   LDS-tiled stencil: each visited cell's cascade of the window
//             (world.hpp:79-153) against the snapshot, with the settling
//             factor of its V events, 1 - (1 - settling)^V; transfers summed
//             in int64, then master += deltas and the next snapshot, in one
//             pass.
//   k_ema     World::erode's EMA flow fields (world.hpp:69-76).
//   small fills/conversions for upload, download and zeroing.
//
// Layout in HBM (all SoA, cell (x,y) at x*ny + y):
//   zm int64 master height, zs[2] float32 snapshot (ping-pong),
//   dz int64 window pickup deltas, vis[2] uint32 cascade events (ping-pong),
//   sm/ss/ds the same for the loose-sediment layer (layers == 2),
//   trk 4 x int64 tracks, ema 4 x float32 fields,
//   particles: P0 float4 (pos.x, pos.y, pos.z, sediment),
//              P1 float4 (speed.x, speed.y, speed.z, age bits),
//              ledger int64 (fixed-point load), lastc int32.
#include "swe_kernels.h"
#include "swe_step.h"

#include "swe_cascade.hpp"
#include "swe_fly.hpp"

namespace swe {

// ---------------------------------------------------------------------------
// Particle march
// ---------------------------------------------------------------------------
// Global-memory stencil: five-point heights of the window's snapshot
template <bool PERIODIC>
struct GlobalMap {
  const float* S;
  const float* SS;
  int32_t cx, cy, nx, ny;
  __device__ float h(int dx, int dy) const {
    int32_t x = cx + dx, y = cy + dy;
    if (PERIODIC) {
      x = x < 0 ? x + nx : (x >= nx ? x - nx : x);
      y = y < 0 ? y + ny : (y >= ny ? y - ny : y);
    }
    return S[(int64_t)x * ny + y];
  }
  __device__ float s0() const { return SS[(int64_t)cx * ny + cy]; }
};

// Device-scope atomics into the window's delta, event and track maps
struct GlobalAcc {
  int64_t* D;
  int64_t* DS;
  uint32_t* V;
  int64_t* T;
  int64_t ncell;
  __device__ void add_pickup(int64_t cell, int64_t q, bool both) {
    atomicAdd((unsigned long long*)&D[cell], (unsigned long long)q);
    if (both) atomicAdd((unsigned long long*)&DS[cell], (unsigned long long)q);
  }
  __device__ void add_event(int64_t cell) { atomicAdd(&V[cell], 1u); }
  __device__ void add_track(int64_t cell, float sed, f3 sp) {
    atomicAdd((unsigned long long*)&T[cell], (unsigned long long)fx_from_f32(sed));
    atomicAdd((unsigned long long*)&T[ncell + cell], (unsigned long long)fx_from_f32(sp.x));
    atomicAdd((unsigned long long*)&T[2 * ncell + cell], (unsigned long long)fx_from_f32(sp.y));
    atomicAdd((unsigned long long*)&T[3 * ncell + cell], (unsigned long long)fx_from_f32(sp.z));
  }
};

template <bool PERIODIC, bool TRACK, int LAYERS, bool DEPOSIT, bool FIRST>
__global__ __launch_bounds__(256) void k_march(MarchArgs a) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  int64_t lost_here = 0;
  uint32_t steps = 0;
  if (i < a.n) {
    FlyConsts k;
    k.nx = a.nx;
    k.ny = a.ny;
    k.rcp_nx = a.rcp_nx;
    k.rcp_ny = a.rcp_ny;
    k.mapscale = a.mapscale;
    k.max_age = a.p.max_age;
    set_boundarylayer(k, a.p.boundarylayer);
    k.suspension = a.p.suspension;
    k.gravity = a.p.gravity;
    k.abrasion = a.p.abrasion;
    k.ps = mk(a.ps[0], a.ps[1], a.ps[2]);
    k.nps = mk(a.nps[0], a.nps[1], a.nps[2]);
    const uint64_t pk = particle_key(cycle_key(a.p.seed, *a.cycle), (uint64_t)i);
    Particle w;
    if (FIRST) {
      spawn(pk, a.nx, a.ny, a.p.spawn, w);
    } else {
      const float4 q0 = a.P0[i];
      const float4 q1 = a.P1[i];
      w.px = q0.x; w.py = q0.y; w.pz = q0.z; w.sed = q0.w;
      w.sx = q1.x; w.sy = q1.y; w.sz = q1.z;
      w.age = (int32_t)f2u(q1.w);
      w.ledger = a.L[i];
      w.lastc = a.LC[i];
    }
    GlobalAcc acc{a.D, a.DS, a.V, a.T, a.ncell};
    for (int32_t s = 0; s < a.R && w.age >= 0; s++) {
      int32_t cx, cy;
      int64_t dead_cell = -1;
      bool die = !cell_of<PERIODIC>(w.px, w.pz, k, cx, cy);  // wind.hpp:30-38
      if (!die) {
        GlobalMap<PERIODIC> m{a.S, a.SS, cx, cy, a.nx, a.ny};
        if (fly<PERIODIC, TRACK, LAYERS>(w, pk, cx, cy, k, m, acc) == kDied) {
          die = true;
          dead_cell = (int64_t)cx * a.ny + cy;
        } else {
          steps++;
        }
      }
      if (die) {
        if (DEPOSIT) {
          if (w.ledger != 0) {
            const int64_t c = dead_cell >= 0 ? dead_cell : (int64_t)w.lastc;
            atomicAdd((unsigned long long*)&a.D[c], (unsigned long long)w.ledger);
            if (LAYERS == 2) atomicAdd((unsigned long long*)&a.DS[c], (unsigned long long)w.ledger);
          }
        } else {
          lost_here += w.ledger;
        }
        w.age = -1;
      }
    }
    a.P0[i] = make_float4(w.px, w.py, w.pz, w.sed);
    a.P1[i] = make_float4(w.sx, w.sy, w.sz, u2f((uint32_t)w.age));
    a.L[i] = w.ledger;
    a.LC[i] = w.lastc;
  }
  if (!DEPOSIT) {
    // carried-off mass: one atomic per wave into one of kLostSlots counters
    for (int off = 32; off > 0; off >>= 1) lost_here += __shfl_xor(lost_here, off, 64);
    if ((threadIdx.x & 63) == 0 && lost_here != 0) {
      const uint32_t slot = (blockIdx.x * 4 + (threadIdx.x >> 6)) & (kLostSlots - 1);
      atomicAdd((unsigned long long*)&a.lost[slot], (unsigned long long)lost_here);
    }
  }
  for (int off = 32; off > 0; off >>= 1) steps += __shfl_xor(steps, off, 64);
  if ((threadIdx.x & 63) == 0 && steps != 0) {
    const uint32_t slot = (blockIdx.x * 4 + (threadIdx.x >> 6)) & (kStatSlots - 1);
    atomicAdd((unsigned long long*)&a.stat[kStatSteps * kStatSlots + slot], (unsigned long long)steps);
  }
}

// ---------------------------------------------------------------------------
// Cascade relax + apply (LDS-tiled stencil)
// ---------------------------------------------------------------------------
// Tile of kTX x kTY cells (y contiguous).  LDS holds the snapshot with a
// 2-cell halo, the event counts with a 1-cell halo, and for every cell of the
// 1-ring region the fixed-point amount each of its 8 neighbours receives
// from that cell's cascade of the window.
constexpr int kTX = 8;
constexpr int kTY = 64;
constexpr int kHX = kTX + 4, kHY = kTY + 4;  // snapshot region
constexpr int kRX = kTX + 2, kRY = kTY + 2;  // ring region
constexpr int kRelaxThreads = 256;

// world.hpp:83-92 neighbour order
constexpr int32_t c_offx[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
constexpr int32_t c_offy[8] = {-1, 0, 1, -1, 1, -1, 0, 1};

template <bool PERIODIC, int LAYERS>
__global__ __launch_bounds__(kRelaxThreads) void k_relax(RelaxArgs a) {
  __shared__ float sh[kHX * kHY];
  __shared__ float shs[LAYERS == 2 ? kHX * kHY : 1];
  __shared__ uint32_t sv[kRX * kRY];
  __shared__ int64_t sx[kRX * kRY * 8];

  const int32_t nx = a.nx, ny = a.ny;
  // XCD-aware tile order: consecutive workgroups land on different XCDs, so
  // hand each XCD a contiguous band of tiles (shares halo lines in its L2).
  // Workgroup b runs on XCD b % 8 as that XCD's (b / 8)-th workgroup; XCD k
  // owns the contiguous tile range starting at k*q + min(k, r).  Bijective
  // for any grid size.
  const uint32_t nb = gridDim.x;
  const uint32_t b = blockIdx.x;
  const uint32_t q = nb / 8, rr = nb % 8, xcd = b % 8;
  const uint32_t t = xcd * q + (xcd < rr ? xcd : rr) + b / 8;
  const int32_t tiles_y = (ny + kTY - 1) / kTY;
  const int32_t x0 = (int32_t)(t / tiles_y) * kTX;
  const int32_t y0 = (int32_t)(t % tiles_y) * kTY;

  // load snapshot (2-halo) and event counts (1-halo)
  for (int e = threadIdx.x; e < kHX * kHY; e += kRelaxThreads) {
    const int32_t lx = e / kHY, ly = e % kHY;
    int32_t gx = x0 + lx - 2, gy = y0 + ly - 2;
    float v = 0.0f, vs = 0.0f;
    if (PERIODIC) {
      gx = wrap(gx, nx, a.rcp_nx);
      gy = wrap(gy, ny, a.rcp_ny);
      v = a.S[(int64_t)gx * ny + gy];
      if (LAYERS == 2) vs = a.SS[(int64_t)gx * ny + gy];
    } else if (gx >= 0 && gy >= 0 && gx < nx && gy < ny) {
      v = a.S[(int64_t)gx * ny + gy];
      if (LAYERS == 2) vs = a.SS[(int64_t)gx * ny + gy];
    }
    sh[e] = v;
    if (LAYERS == 2) shs[e] = vs;
  }
  for (int e = threadIdx.x; e < kRX * kRY; e += kRelaxThreads) {
    const int32_t lx = e / kRY, ly = e % kRY;
    int32_t gx = x0 + lx - 1, gy = y0 + ly - 1;
    uint32_t v = 0;
    if (PERIODIC) {
      gx = wrap(gx, nx, a.rcp_nx);
      gy = wrap(gy, ny, a.rcp_ny);
      v = a.V[(int64_t)gx * ny + gy];
    } else if (gx >= 0 && gy >= 0 && gx < nx && gy < ny) {
      v = a.V[(int64_t)gx * ny + gy];
    }
    sv[e] = v;
  }
  __syncthreads();

  // cascade events of every ring cell with visits
  for (int e = threadIdx.x; e < kRX * kRY; e += kRelaxThreads) {
    const uint32_t v = sv[e];
    int64_t* out = &sx[e * 8];
    if (v == 0) {
#pragma unroll
      for (int k = 0; k < 8; k++) out[k] = 0;
      continue;
    }
    const int32_t lx = e / kRY, ly = e % kRY;  // ring coords; snapshot coords +1
    const int32_t gx = x0 + lx - 1, gy = y0 + ly - 1;
    float hn[8];
    uint32_t valid = 0;
    const int32_t ic = (lx + 1) * kHY + ly + 1;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      hn[k] = sh[ic + c_offx[k] * kHY + c_offy[k]];
      bool ok = true;
      if (!PERIODIC) {
        const int32_t qx = gx + c_offx[k], qy = gy + c_offy[k];
        ok = qx >= 0 && qy >= 0 && qx < nx && qy < ny;
      }
      valid |= (ok ? 1u : 0u) << k;
      out[k] = 0;
    }
    const float hc = sh[ic];
    const float sc = LAYERS == 2 ? shs[ic] : 0.0f;
    const auto sed_of = [&](int32_t k) { return shs[ic + off_x(k) * kHY + off_y(k)]; };
    cascade_cell_dev<LAYERS>(hn, sed_of, hc, sc, valid, settle_factor_dev(a.settling, v), a.maxdiff,
                             [&](int32_t k, int64_t amt) { out[k] = amt; });
  }
  __syncthreads();

  // gather: own events' centre share + what each neighbour's events sent here
  for (int e = threadIdx.x; e < kTX * kTY; e += kRelaxThreads) {
    const int32_t lx = e / kTY, ly = e % kTY;
    const int32_t gx = x0 + lx, gy = y0 + ly;
    if (gx >= nx || gy >= ny) continue;
    const int32_t rc = (lx + 1) * kRY + (ly + 1);
    int64_t cas = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) cas -= sx[rc * 8 + k];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      // neighbour y = x - off_k sends through its offset k to x
      const int32_t ry = rc - c_offx[k] * kRY - c_offy[k];
      cas += sx[ry * 8 + k];
    }
    const int64_t c = (int64_t)gx * ny + gy;
    const int64_t z = a.zm_in[c] + a.dz[c] + cas;
    a.zm_out[c] = z;
    a.zs_out[c] = f32_from_fx(z);
    a.dz[c] = 0;
    a.v_clear[c] = 0;
    if (LAYERS == 2) {
      int64_t s = a.sm_in[c] + a.ds[c] + cas;
      if (s < 0) s = 0;
      a.sm_out[c] = s;
      a.ss_out[c] = f32_from_fx(s);
      a.ds[c] = 0;
    }
  }
}

// ---------------------------------------------------------------------------
// Colour-ordered relax: the window's cascades one colour class after another
// ---------------------------------------------------------------------------
// Masters after the window's pickups, their snapshots, and the next window's
// event buffer cleared.
template <int LAYERS>
__global__ void k_colour_begin(RelaxArgs a, int64_t n) {
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    const int64_t z = a.zm_in[c] + a.dz[c];
    a.zm_out[c] = z;
    a.zs_out[c] = f32_from_fx(z);
    a.dz[c] = 0;
    a.v_clear[c] = 0;
    if (LAYERS == 2) {
      const int64_t v = a.sm_in[c] + a.ds[c];
      a.sm_out[c] = v;
      a.ss_out[c] = f32_from_fx(v);
      a.ds[c] = 0;
    }
  }
}

// One colour class (cx, cy): every cell of the class with events runs its
// cascade against the current map (zs_out / ss_out) and adds its transfers
// to the masters at once.  Cells of one class are at least three apart, so
// no two threads touch the same cell: plain stores.  All loads of a cell's
// 3x3 block (snapshot and masters) are issued together and its nine
// masters and snapshots written once, so a class costs two memory round
// trips.  On maps under three cells along an axis the block's cells alias
// (the periodic wrap reaches the centre or a neighbour twice): there the
// transfers are added to memory one after another.
template <bool PERIODIC>
__device__ __forceinline__ int64_t colour_neighbour(int32_t x, int32_t y, int k, int32_t nx, int32_t ny, bool& ok) {
  int32_t qx = x + c_offx[k], qy = y + c_offy[k];
  ok = true;
  if (PERIODIC) {
    qx = qx < 0 ? qx + nx : (qx >= nx ? qx - nx : qx);
    qy = qy < 0 ? qy + ny : (qy >= ny ? qy - ny : qy);
  } else {
    ok = qx >= 0 && qy >= 0 && qx < nx && qy < ny;
  }
  return (int64_t)qx * ny + qy;
}

// cell (x, y) of a map of nx x ny cells starting at cell `base` of the
// context's arrays
template <bool PERIODIC, int LAYERS>
__device__ __forceinline__ void colour_relax_cell(const RelaxArgs& a, int64_t base, int32_t nx, int32_t ny, int32_t x,
                                                  int32_t y) {
  const int64_t c = base + (int64_t)x * ny + y;
  const uint32_t v = a.V[c];
  if (v == 0) return;
  int64_t zn[8], snm[8];
  float hn[8], sn[8];
  uint32_t valid = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    bool ok;
    const int64_t q = base + colour_neighbour<PERIODIC>(x, y, k, nx, ny, ok);
    hn[k] = ok ? a.zs_out[q] : 0.0f;
    zn[k] = ok ? a.zm_out[q] : 0;
    sn[k] = (LAYERS == 2 && ok) ? a.ss_out[q] : 0.0f;
    snm[k] = (LAYERS == 2 && ok) ? a.sm_out[q] : 0;
    valid |= (ok ? 1u : 0u) << k;
  }
  const float hc = a.zs_out[c];
  const float sc = LAYERS == 2 ? a.ss_out[c] : 0.0f;
  const int64_t zc = a.zm_out[c];
  const int64_t scm = LAYERS == 2 ? a.sm_out[c] : 0;
  const bool alias = nx < 3 || ny < 3;
  int64_t d[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int64_t centre = 0;
  const auto sed_of = [&](int32_t k) { return sn[k]; };
  // the neighbour index k of a transfer is data-dependent: select into the
  // eight accumulators (an indexed register array would live in scratch)
  cascade_cell_dev<LAYERS>(hn, sed_of, hc, sc, valid, settle_factor_dev(a.settling, v), a.maxdiff,
                           [&](int32_t k, int64_t amt) {
#pragma unroll
                             for (int j = 0; j < 8; j++) d[j] += k == j ? amt : 0;
                             centre -= amt;
                           });
  if (alias) {
#pragma unroll
    for (int k = 0; k < 8; k++) {
      bool ok;
      const int64_t q = base + colour_neighbour<PERIODIC>(x, y, k, nx, ny, ok);
      if (!ok || d[k] == 0) continue;
      a.zm_out[q] += d[k];
      if (LAYERS == 2) a.sm_out[q] += d[k];
    }
    a.zm_out[c] += centre;
    if (LAYERS == 2) a.sm_out[c] += centre;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      bool ok;
      const int64_t q = base + colour_neighbour<PERIODIC>(x, y, k, nx, ny, ok);
      if (!ok) continue;
      a.zs_out[q] = f32_from_fx(a.zm_out[q]);
      if (LAYERS == 2) a.ss_out[q] = f32_from_fx(a.sm_out[q]);
    }
    a.zs_out[c] = f32_from_fx(a.zm_out[c]);
    if (LAYERS == 2) a.ss_out[c] = f32_from_fx(a.sm_out[c]);
    return;
  }
#pragma unroll
  for (int k = 0; k < 8; k++) {
    if (d[k] == 0) continue;
    bool ok;
    const int64_t q = base + colour_neighbour<PERIODIC>(x, y, k, nx, ny, ok);
    const int64_t z = zn[k] + d[k];
    a.zm_out[q] = z;
    a.zs_out[q] = f32_from_fx(z);
    if (LAYERS == 2) {
      const int64_t sm = snm[k] + d[k];
      a.sm_out[q] = sm;
      a.ss_out[q] = f32_from_fx(sm);
    }
  }
  if (centre != 0) {
    a.zm_out[c] = zc + centre;
    a.zs_out[c] = f32_from_fx(zc + centre);
    if (LAYERS == 2) {
      a.sm_out[c] = scm + centre;
      a.ss_out[c] = f32_from_fx(scm + centre);
    }
  }
}

template <bool PERIODIC, int LAYERS>
__global__ __launch_bounds__(256) void k_colour_pass(RelaxArgs a, int32_t cx, int32_t cy, int32_t ni, int32_t nj) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)ni * nj) return;
  const int32_t x = colour_cell(a.nx, PERIODIC, cx, (int32_t)(t / nj));
  const int32_t y = colour_cell(a.ny, PERIODIC, cy, (int32_t)(t % nj));
  colour_relax_cell<PERIODIC, LAYERS>(a, 0, a.nx, a.ny, x, y);
}

// The same class (cx, cy) on every terrain of a batch at once (terrains
// share no cell, so their classes run side by side; a terrain without the
// class has none of the launch's cells).  pre[k] = the class's cells on
// terrains 0 .. k-1; thread t takes the terrain with pre[k] <= t < pre[k+1].
template <bool PERIODIC, int LAYERS>
__global__ __launch_bounds__(256) void k_colour_pass_batch(RelaxArgs a, const Terrain* __restrict__ terr,
                                                           const int64_t* __restrict__ pre, int32_t nterr, int32_t cx,
                                                           int32_t cy) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= pre[nterr]) return;
  int32_t lo = 0, hi = nterr - 1;
  while (lo < hi) {
    const int32_t mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  const int64_t base = terr[lo].cell0;
  const int32_t nx = terr[lo].nx, ny = terr[lo].ny;
  const int64_t u = t - pre[lo];
  const int32_t nj = colour_count(ny, PERIODIC, cy);
  const int32_t x = colour_cell(nx, PERIODIC, cx, (int32_t)(u / nj));
  const int32_t y = colour_cell(ny, PERIODIC, cy, (int32_t)(u % nj));
  colour_relax_cell<PERIODIC, LAYERS>(a, base, nx, ny, x, y);
}

// the sediment layer clamped at zero after the last class
__global__ void k_colour_end(RelaxArgs a, int64_t n) {
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = a.sm_out[c];
    if (v < 0) {
      a.sm_out[c] = 0;
      a.ss_out[c] = 0.0f;
    }
  }
}

// ---------------------------------------------------------------------------
// Elementwise kernels
// ---------------------------------------------------------------------------
__global__ void k_zero64(int64_t* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}
__global__ void k_zero32(uint32_t* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}
// float -> fixed-point master + snapshot
__global__ void k_from_f32(const float* in, int64_t* master, float* snap, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t m = fx_from_f32(in[i]);
    master[i] = m;
    snap[i] = f32_from_fx(m);
  }
}
__global__ void k_snap(const int64_t* master, float* snap, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    snap[i] = f32_from_fx(master[i]);
}
// world.hpp:69-76: f = (1-lrate)*f + lrate*track, four fields
__global__ void k_ema(float* ema, const int64_t* trk, int64_t n4, float lrate) {
  const float a = 1.0f - lrate;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (int64_t)gridDim.x * blockDim.x)
    ema[i] = a * ema[i] + lrate * f32_from_fx(trk[i]);
}
// updatenode (cellpool.h:302-333): per cell the vertex position P =
// (x, mapscale*h, y), the normal, T - P and B - P with T, B the +x / +y
// neighbours (height 0 off a bounded map, wrapped on a periodic one)
template <bool PERIODIC>
__global__ void k_mesh(const float* S, float* out, int32_t nx, int32_t ny, float mapscale) {
  const int64_t n = (int64_t)nx * ny;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    const int32_t cx = (int32_t)(c / ny), cy = (int32_t)(c % ny);
    const int64_t row = (int64_t)cx * ny;
    const float h = S[c];
    float hyp = 0.0f, hym = 0.0f, hxp = 0.0f, hxm = 0.0f;
    bool g1, g2, g3, g4;
    if (PERIODIC) {
      const int32_t xp = cx + 1 == nx ? 0 : cx + 1, xm = cx == 0 ? nx - 1 : cx - 1;
      const int32_t yp = cy + 1 == ny ? 0 : cy + 1, ym = cy == 0 ? ny - 1 : cy - 1;
      hyp = S[row + yp];
      hym = S[row + ym];
      hxp = S[(int64_t)xp * ny + cy];
      hxm = S[(int64_t)xm * ny + cy];
      g1 = g2 = g3 = g4 = true;
    } else {
      const bool bxp = cx + 1 < nx, bxm = cx > 0, byp = cy + 1 < ny, bym = cy > 0;
      g1 = bxp && byp;
      g2 = bxm && bym;
      g3 = bxp && bym;
      g4 = bxm && byp;
      if (byp) hyp = S[row + cy + 1];
      if (bym) hym = S[row + cy - 1];
      if (bxp) hxp = S[row + ny + cy];
      if (bxm) hxm = S[row - ny + cy];
    }
    const f3 nn = surface_normal(h, hyp, hxp, hym, hxm, g1, g2, g3, g4, mapscale);
    const float px = (float)cx, pz = (float)cy;
    const f3 P = mk(px, mapscale * h, pz);
    const f3 T = mk((float)(cx + 1), mapscale * hxp, pz);
    const f3 B = mk(px, mapscale * hyp, (float)(cy + 1));
    const f3 t = sub(T, P), b = sub(B, P);
    float* o = out + 12 * c;
    o[0] = P.x; o[1] = P.y; o[2] = P.z;
    o[3] = nn.x; o[4] = nn.y; o[5] = nn.z;
    o[6] = t.x; o[7] = t.y; o[8] = t.z;
    o[9] = b.x; o[10] = b.y; o[11] = b.z;
  }
}

// SimpleWindErosion.cpp:224-230: colour = 0.5 (1 + erf(2 m)) per momentum axis;
// the x, y, z planes of n cells lie `plane` floats apart
__global__ void k_windmap(const float* mom, int64_t plane, float* out, int64_t n) {
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    const float mx = mom[c], my = mom[plane + c], mz = mom[2 * plane + c];
    out[4 * c + 0] = 0.5f * (1.0f + erff(2.0f * mx));
    out[4 * c + 1] = 0.5f * (1.0f + erff(2.0f * my));
    out[4 * c + 2] = 0.5f * (1.0f + erff(2.0f * mz));
    out[4 * c + 3] = 1.0f;
  }
}

__global__ void k_bump(uint64_t* cycle) {
  if (threadIdx.x == 0 && blockIdx.x == 0) cycle[0] = cycle[0] + 1;
}
// quad::_normal for every cell of the snapshot (cellpool.h:200-223,478-480)
template <bool PERIODIC>
__global__ void k_normals(const float* S, float* out, int32_t nx, int32_t ny, float mapscale) {
  const int64_t n = (int64_t)nx * ny;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    const int32_t cx = (int32_t)(c / ny), cy = (int32_t)(c % ny);
    const int64_t row = (int64_t)cx * ny;
    const float h = S[c];
    float hyp = 0.0f, hym = 0.0f, hxp = 0.0f, hxm = 0.0f;
    bool g1, g2, g3, g4;
    if (PERIODIC) {
      const int32_t xp = cx + 1 == nx ? 0 : cx + 1, xm = cx == 0 ? nx - 1 : cx - 1;
      const int32_t yp = cy + 1 == ny ? 0 : cy + 1, ym = cy == 0 ? ny - 1 : cy - 1;
      hyp = S[row + yp];
      hym = S[row + ym];
      hxp = S[(int64_t)xp * ny + cy];
      hxm = S[(int64_t)xm * ny + cy];
      g1 = g2 = g3 = g4 = true;
    } else {
      const bool bxp = cx + 1 < nx, bxm = cx > 0, byp = cy + 1 < ny, bym = cy > 0;
      g1 = bxp && byp;
      g2 = bxm && bym;
      g3 = bxp && bym;
      g4 = bxm && byp;
      if (byp) hyp = S[row + cy + 1];
      if (bym) hym = S[row + cy - 1];
      if (bxp) hxp = S[row + ny + cy];
      if (bxm) hxm = S[row - ny + cy];
    }
    const f3 nn = surface_normal(h, hyp, hxp, hym, hxm, g1, g2, g3, g4, mapscale);
    out[c] = nn.x;
    out[n + c] = nn.y;
    out[2 * n + c] = nn.z;
  }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static inline unsigned grid_for(int64_t n, int threads) {
  int64_t g = (n + threads - 1) / threads;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (unsigned)g;
}

template <bool P, bool T, int L, bool D>
static void march_first(const MarchArgs& a, bool first, hipStream_t s) {
  const unsigned g = (unsigned)((a.n + 255) / 256);
  if (first)
    hipLaunchKernelGGL((k_march<P, T, L, D, true>), dim3(g), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((k_march<P, T, L, D, false>), dim3(g), dim3(256), 0, s, a);
}
template <bool P, bool T, int L>
static void march_dep(const MarchArgs& a, bool first, bool dep, hipStream_t s) {
  if (dep) march_first<P, T, L, true>(a, first, s);
  else march_first<P, T, L, false>(a, first, s);
}
template <bool P, bool T>
static void march_lay(const MarchArgs& a, bool first, bool dep, int layers, hipStream_t s) {
  if (layers == 2) march_dep<P, T, 2>(a, first, dep, s);
  else march_dep<P, T, 1>(a, first, dep, s);
}
template <bool P>
static void march_trk(const MarchArgs& a, bool first, bool dep, bool trk, int layers, hipStream_t s) {
  if (trk) march_lay<P, true>(a, first, dep, layers, s);
  else march_lay<P, false>(a, first, dep, layers, s);
}

void launch_march(const MarchArgs& a, bool first, hipStream_t s) {
  if (a.n <= 0) return;
  const bool dep = a.p.deposit_on_death != 0, trk = a.p.track != 0;
  if (a.p.periodic) march_trk<true>(a, first, dep, trk, a.p.layers, s);
  else march_trk<false>(a, first, dep, trk, a.p.layers, s);
}

void launch_relax(const RelaxArgs& a, bool periodic, int layers, hipStream_t s) {
  const unsigned tiles = (unsigned)(((a.nx + kTX - 1) / kTX) * ((a.ny + kTY - 1) / kTY));
  if (periodic) {
    if (layers == 2) hipLaunchKernelGGL((k_relax<true, 2>), dim3(tiles), dim3(kRelaxThreads), 0, s, a);
    else hipLaunchKernelGGL((k_relax<true, 1>), dim3(tiles), dim3(kRelaxThreads), 0, s, a);
  } else {
    if (layers == 2) hipLaunchKernelGGL((k_relax<false, 2>), dim3(tiles), dim3(kRelaxThreads), 0, s, a);
    else hipLaunchKernelGGL((k_relax<false, 1>), dim3(tiles), dim3(kRelaxThreads), 0, s, a);
  }
}

void launch_relax_colour(const RelaxArgs& a, bool periodic, int layers, bool begin, hipStream_t s) {
  const int64_t n = (int64_t)a.nx * a.ny;
  if (begin) {
    if (layers == 2) hipLaunchKernelGGL(k_colour_begin<2>, dim3(grid_for(n, 256)), dim3(256), 0, s, a, n);
    else hipLaunchKernelGGL(k_colour_begin<1>, dim3(grid_for(n, 256)), dim3(256), 0, s, a, n);
  }
  const int32_t ncx = colour_classes(a.nx, periodic), ncy = colour_classes(a.ny, periodic);
  for (int32_t cx = 0; cx < ncx; cx++)
    for (int32_t cy = 0; cy < ncy; cy++) {
      const int32_t ni = colour_count(a.nx, periodic, cx), nj = colour_count(a.ny, periodic, cy);
      const int64_t m = (int64_t)ni * nj;
      if (m == 0) continue;
      const dim3 g((unsigned)((m + 255) / 256));
      if (periodic) {
        if (layers == 2) hipLaunchKernelGGL((k_colour_pass<true, 2>), g, dim3(256), 0, s, a, cx, cy, ni, nj);
        else hipLaunchKernelGGL((k_colour_pass<true, 1>), g, dim3(256), 0, s, a, cx, cy, ni, nj);
      } else {
        if (layers == 2) hipLaunchKernelGGL((k_colour_pass<false, 2>), g, dim3(256), 0, s, a, cx, cy, ni, nj);
        else hipLaunchKernelGGL((k_colour_pass<false, 1>), g, dim3(256), 0, s, a, cx, cy, ni, nj);
      }
    }
  if (layers == 2) hipLaunchKernelGGL(k_colour_end, dim3(grid_for(n, 256)), dim3(256), 0, s, a, n);
}

void launch_relax_colour_batch(const RelaxArgs& a, const Terrain* terr, const int64_t* pre_dev,
                               const int64_t* pre_host, int32_t nterr, int64_t ncell, bool periodic, int layers,
                               hipStream_t s) {
  for (int32_t cx = 0; cx < kColourAxisMax; cx++)
    for (int32_t cy = 0; cy < kColourAxisMax; cy++) {
      const int32_t cls = cx * kColourAxisMax + cy;
      const int64_t m = pre_host[(int64_t)cls * (nterr + 1) + nterr];
      if (m == 0) continue;
      const int64_t* pre = pre_dev + (int64_t)cls * (nterr + 1);
      const dim3 g((unsigned)((m + 255) / 256));
      if (periodic) {
        if (layers == 2) hipLaunchKernelGGL((k_colour_pass_batch<true, 2>), g, dim3(256), 0, s, a, terr, pre, nterr, cx, cy);
        else hipLaunchKernelGGL((k_colour_pass_batch<true, 1>), g, dim3(256), 0, s, a, terr, pre, nterr, cx, cy);
      } else {
        if (layers == 2) hipLaunchKernelGGL((k_colour_pass_batch<false, 2>), g, dim3(256), 0, s, a, terr, pre, nterr, cx, cy);
        else hipLaunchKernelGGL((k_colour_pass_batch<false, 1>), g, dim3(256), 0, s, a, terr, pre, nterr, cx, cy);
      }
    }
  if (layers == 2) hipLaunchKernelGGL(k_colour_end, dim3(grid_for(ncell, 256)), dim3(256), 0, s, a, ncell);
}

void colour_prefix(const int32_t* nx, const int32_t* ny, int32_t nterr, bool periodic, int64_t* pre) {
  for (int32_t cx = 0; cx < kColourAxisMax; cx++)
    for (int32_t cy = 0; cy < kColourAxisMax; cy++) {
      int64_t* p = pre + (int64_t)(cx * kColourAxisMax + cy) * (nterr + 1);
      p[0] = 0;
      for (int32_t k = 0; k < nterr; k++) {
        const int64_t n = (cx < colour_classes(nx[k], periodic) && cy < colour_classes(ny[k], periodic))
                              ? (int64_t)colour_count(nx[k], periodic, cx) * colour_count(ny[k], periodic, cy)
                              : 0;
        p[k + 1] = p[k] + n;
      }
    }
}

void launch_zero64(int64_t* p, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_zero64, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n);
}
void launch_zero32(uint32_t* p, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_zero32, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n);
}
void launch_from_f32(const float* in, int64_t* master, float* snap, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_from_f32, dim3(grid_for(n, 256)), dim3(256), 0, s, in, master, snap, n);
}
void launch_snap(const int64_t* master, float* snap, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_snap, dim3(grid_for(n, 256)), dim3(256), 0, s, master, snap, n);
}
void launch_ema(float* ema, const int64_t* trk, int64_t n4, float lrate, hipStream_t s) {
  if (n4 > 0) hipLaunchKernelGGL(k_ema, dim3(grid_for(n4, 256)), dim3(256), 0, s, ema, trk, n4, lrate);
}
void launch_mesh(const float* S, float* out, int32_t nx, int32_t ny, float mapscale, bool periodic, hipStream_t s) {
  const int64_t n = (int64_t)nx * ny;
  if (periodic) hipLaunchKernelGGL((k_mesh<true>), dim3(grid_for(n, 256)), dim3(256), 0, s, S, out, nx, ny, mapscale);
  else hipLaunchKernelGGL((k_mesh<false>), dim3(grid_for(n, 256)), dim3(256), 0, s, S, out, nx, ny, mapscale);
}
void launch_windmap(const float* mom, int64_t plane, float* out, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_windmap, dim3(grid_for(n, 256)), dim3(256), 0, s, mom, plane, out, n);
}
void launch_bump(uint64_t* cycle, hipStream_t s) {
  hipLaunchKernelGGL(k_bump, dim3(1), dim3(64), 0, s, cycle);
}
void launch_normals(const float* S, float* out, int32_t nx, int32_t ny, float mapscale, bool periodic,
                    hipStream_t s) {
  const int64_t n = (int64_t)nx * ny;
  if (periodic) hipLaunchKernelGGL((k_normals<true>), dim3(grid_for(n, 256)), dim3(256), 0, s, S, out, nx, ny, mapscale);
  else hipLaunchKernelGGL((k_normals<false>), dim3(grid_for(n, 256)), dim3(256), 0, s, S, out, nx, ny, mapscale);
}

// ---------------------------------------------------------------------------
// Known-answer entry points
// ---------------------------------------------------------------------------
// fly()'s writes discarded: the call's pickup is the particle's ledger
struct NullAcc {
  __device__ void add_pickup(int64_t, int64_t, bool) {}
  __device__ void add_event(int64_t) {}
  __device__ void add_track(int64_t, float, f3) {}
};

template <bool PERIODIC, int LAYERS>
__global__ __launch_bounds__(64) void k_debug_fly(DebugFlyArgs a) {
  const int32_t i = (int32_t)(blockIdx.x * 64 + threadIdx.x);
  if (i >= a.n) return;
  FlyConsts k;
  k.nx = a.nx;
  k.ny = a.ny;
  k.rcp_nx = a.rcp_nx;
  k.rcp_ny = a.rcp_ny;
  k.mapscale = a.mapscale;
  k.max_age = a.p.max_age;
  set_boundarylayer(k, a.p.boundarylayer);
  k.suspension = a.p.suspension;
  k.gravity = a.p.gravity;
  k.abrasion = a.p.abrasion;
  k.ps = mk(a.ps[0], a.ps[1], a.ps[2]);
  k.nps = mk(a.nps[0], a.nps[1], a.nps[2]);
  float* st = a.st + (int64_t)7 * i;
  Particle w;
  w.px = st[0];
  w.py = st[1];
  w.pz = st[2];
  w.sx = st[3];
  w.sy = st[4];
  w.sz = st[5];
  w.sed = st[6];
  w.age = a.age[i];
  w.ledger = 0;
  w.lastc = 0;
  int32_t cx, cy;
  bool ok = cell_of<PERIODIC>(w.px, w.pz, k, cx, cy);  // wind.hpp:30-38
  if (ok) {
    GlobalMap<PERIODIC> m{a.S, a.SS, cx, cy, a.nx, a.ny};
    NullAcc acc;
    ok = fly<PERIODIC, false, LAYERS>(w, a.key[i], cx, cy, k, m, acc) == kStepped;
  }
  if (!ok) w.age = -1;  // the march's mark of a dead particle
  st[0] = w.px;
  st[1] = w.py;
  st[2] = w.pz;
  st[3] = w.sx;
  st[4] = w.sy;
  st[5] = w.sz;
  st[6] = w.sed;
  a.age[i] = w.age;
  a.alive[i] = ok ? 1 : 0;
  a.pick[i] = w.ledger;
}

template <bool PERIODIC, int LAYERS>
__global__ void k_debug_cascade(DebugCascadeArgs a) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int64_t c = (int64_t)a.x * a.ny + a.y;
  float hn[8], sn[8];
  int64_t q[8];
  uint32_t valid = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    bool ok;
    q[k] = colour_neighbour<PERIODIC>(a.x, a.y, k, a.nx, a.ny, ok);
    hn[k] = ok ? a.S[q[k]] : 0.0f;
    sn[k] = (LAYERS == 2 && ok) ? a.SS[q[k]] : 0.0f;
    valid |= (ok ? 1u : 0u) << k;
  }
  const float hc = a.S[c];
  const float sc = LAYERS == 2 ? a.SS[c] : 0.0f;
  int64_t d[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const auto sed_of = [&](int32_t k) { return sn[k]; };
  cascade_cell_dev<LAYERS>(hn, sed_of, hc, sc, valid, settle_factor_dev(a.p.settling, a.events), a.p.maxdiff,
                           [&](int32_t k, int64_t amount) {
#pragma unroll
                             for (int j = 0; j < 8; j++)
                               if (j == k) d[j] += amount;
                           });
  int32_t n = 0;
  for (int k = 0; k < 8; k++) {
    if (!((valid >> k) & 1u)) continue;
    a.cell[n] = q[k];
    a.fx[n] = d[k];
    n++;
  }
  *a.count = n;
}

void launch_debug_fly(const DebugFlyArgs& a, hipStream_t s) {
  if (a.n <= 0) return;
  const dim3 g((unsigned)((a.n + 63) / 64));
  if (a.p.periodic) {
    if (a.p.layers == 2) hipLaunchKernelGGL((k_debug_fly<true, 2>), g, dim3(64), 0, s, a);
    else hipLaunchKernelGGL((k_debug_fly<true, 1>), g, dim3(64), 0, s, a);
  } else {
    if (a.p.layers == 2) hipLaunchKernelGGL((k_debug_fly<false, 2>), g, dim3(64), 0, s, a);
    else hipLaunchKernelGGL((k_debug_fly<false, 1>), g, dim3(64), 0, s, a);
  }
}

void launch_debug_cascade(const DebugCascadeArgs& a, hipStream_t s) {
  if (a.p.periodic) {
    if (a.p.layers == 2) hipLaunchKernelGGL((k_debug_cascade<true, 2>), dim3(1), dim3(64), 0, s, a);
    else hipLaunchKernelGGL((k_debug_cascade<true, 1>), dim3(1), dim3(64), 0, s, a);
  } else {
    if (a.p.layers == 2) hipLaunchKernelGGL((k_debug_cascade<false, 2>), dim3(1), dim3(64), 0, s, a);
    else hipLaunchKernelGGL((k_debug_cascade<false, 1>), dim3(1), dim3(64), 0, s, a);
  }
}

}  // namespace swe
