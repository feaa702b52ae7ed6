// swe_capi.hip -- the C-ABI of libswe.so (declared in include/swe.h).
//
// A context owns one map on one device: the fixed-point master layers, the
// snapshot/event ping-pong buffers, the particle state and a HIP stream.
// swe_erode() drives World::erode's cycle (world.hpp:36-77) as a sequence of
// snapshot windows: k_march (R fly() calls per particle) then k_relax
// (cascade + apply + next snapshot), then the EMA fields.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <array>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/swe.h"
#include "swe_kernels.h"
#include "swe_step.h"

namespace {

// Staging of swe_normals_noalloc: three float planes of at most
// kNoAllocCells cells.  Every device the library runs on holds its own copy
// of this array until the process ends.  Host code never passes the array's
// name to a kernel or a copy: it uses the device address that
// hipGetSymbolAddress returns for the current device (scratch_of below).
constexpr int64_t kNoAllocCells = (int64_t)1 << 20;
__device__ __align__(256) float scratch[3 * kNoAllocCells];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};

// Resolves this device's copy of `scratch` once per device, with `device`
// current.  Called when a context is created, so the no-allocation call
// itself only reads the table.
hipError_t resolve_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return hipSuccess;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_ptr[device] = (float*)ptr;
  return hipSuccess;
}

float* scratch_of(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}

}  // namespace

struct swe_ctx {
  int device = 0;
  int32_t nx = 0, ny = 0;
  int64_t ncell = 0;
  hipStream_t stream = nullptr;
  swe_params p{};
  std::string err;
  int32_t path = 0;  // 0 auto, 1 global-atomic march, 2 tile-owned step
  int32_t tile_order = 0;  // SWE_TUNE_TILE_ORDER
  int32_t use_graph = 1;   // SWE_TUNE_GRAPH
  int32_t heavy_first = -1;  // SWE_TUNE_HEAVY_FIRST; -1 = default (2 up to 4096 tiles, else 1)
  int32_t shape = 0;         // SWE_TUNE_SHAPE: index into swe::kStepShape
  // one captured map-cycle of the tile path per starting window parity, and
  // the launch configuration it was captured for
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  std::string gkey[2];
  // bumped by every device allocation of the context: a captured graph holds
  // buffer addresses, and a freed buffer's address can come back from
  // hipMalloc for another one, so the generation is part of the graph key
  uint64_t alloc_gen = 0;

  // map (device); index [g & 1] is current, g counts snapshot windows
  uint64_t g = 0;
  int64_t* zmv[2] = {nullptr, nullptr};
  float* zs[2] = {nullptr, nullptr};
  int64_t* smv[2] = {nullptr, nullptr};
  float* ss[2] = {nullptr, nullptr};
  int64_t* trk = nullptr;
  float* ema = nullptr;

  // global-atomic march path
  int64_t* dz = nullptr;
  int64_t* ds = nullptr;
  uint32_t* vis[2] = {nullptr, nullptr};
  int64_t cap = 0;
  float4* P0 = nullptr;
  float4* P1 = nullptr;
  int64_t* L = nullptr;
  int32_t* LC = nullptr;

  // terrains packed into the buffers (a plain map is one terrain)
  int32_t nterr = 1;
  std::vector<int32_t> tnx, tny, ttx, tty, ttile0;
  std::vector<int64_t> tcell0;
  std::vector<uint64_t> tseed;
  std::vector<std::array<float, 3>> tps;
  std::vector<char> tseed_set, tps_set;
  int32_t ntiles = 0;
  swe::Terrain* dterr = nullptr;
  // colour classes of a batch's terrains (swe::colour_prefix), periodic
  // table then bounded, on the device and the host
  int64_t* dcolour = nullptr;
  std::vector<int64_t> hcolour;
  int32_t* dtile_terrain = nullptr;

  // tile-owned step path
  uint32_t* fV[2] = {nullptr, nullptr};
  int64_t fcap = 0;
  swe::PRec* PB[2] = {};  // particle records, ping-pong by window parity
  swe::PRec* FB[3] = {};  // far lists
  uint32_t* farcnt = nullptr;  // 3
  int64_t* sbase[2] = {nullptr, nullptr};   // 9 x tiles: routing sub-range starts
  uint32_t* scnt[2] = {nullptr, nullptr};   // 9 x tiles: routing sub-range lengths
  uint32_t* nin = nullptr;     // 3 x tiles
  uint32_t* csum = nullptr;    // 3 x chunks of kStepChunk tiles
  uint32_t* heavy[2] = {nullptr, nullptr};  // kHeavySlotsMax crowded + light_cap light tiles listed for a window
  uint32_t* hcnt = nullptr;    // 3 x 16: their counts (heavy per XCD, light per XCD)
  uint32_t* hstamp = nullptr;  // per tile: window + 1 when listed
  int32_t* sp_tile = nullptr;
  uint32_t* sp_rank = nullptr;

  // counters (device)
  int64_t* lost = nullptr;     // swe::kLostSlots per terrain
  uint64_t* dstat = nullptr;   // swe::kStatKinds x swe::kStatSlots device statistics
  uint64_t* dcycle = nullptr;  // RNG cycle counter
  uint64_t cycle = 0;          // host mirror

  // timing
  int32_t timing = 0;     // 0 off, n >= 1: every n-th kernel launch timed
  uint64_t timing_tick = 0;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  struct Pending {
    hipEvent_t a, b;
    int kind;
  };
  std::vector<Pending> pending;
  swe_stats_t st{};

  int cur() const { return (int)(g & 1u); }
};

namespace {

int fail(swe_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(c, expr)                                                             \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess)                                                           \
      return fail((c), SWE_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
int dalloc(swe_ctx* c, T** ptr, int64_t count) {
  *ptr = nullptr;
  if (count <= 0) return 0;
  c->alloc_gen++;
  hipError_t e = hipMalloc((void**)ptr, sizeof(T) * (size_t)count);
  if (e != hipSuccess) {
    *ptr = nullptr;
    return fail(c, SWE_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  e = hipMemsetAsync(*ptr, 0, sizeof(T) * (size_t)count, c->stream);
  if (e != hipSuccess) return fail(c, SWE_EHIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
  return 0;
}

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

// glm::normalize on the host, same operation order as the device
void normalize3(const float v[3], float out[3]) {
  const float d = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
  const float r = 1.0f / std::sqrt(d);
  out[0] = v[0] * r;
  out[1] = v[1] * r;
  out[2] = v[2] * r;
}

int check_params(swe_ctx* c, const swe_params* p) {
  if (p->layers != 1 && p->layers != 2) return fail(c, SWE_EINVAL, "layers must be 1 or 2");
  if (p->window < 0) return fail(c, SWE_EINVAL, "window must be >= 0");
  if (p->spawn != 0 && p->spawn != 1) return fail(c, SWE_EINVAL, "spawn must be 0 (uniform) or 1 (edges)");
  if (p->max_age < -1 || p->max_age > (1 << 28)) return fail(c, SWE_EINVAL, "max_age out of range");
  if (p->periodic != 0 && p->periodic != 1) return fail(c, SWE_EINVAL, "periodic must be 0 or 1");
  if (p->track != 0 && p->track != 1) return fail(c, SWE_EINVAL, "track must be 0 or 1");
  if (p->deposit_on_death != 0 && p->deposit_on_death != 1)
    return fail(c, SWE_EINVAL, "deposit_on_death must be 0 or 1");
  if (p->cascade_order < -1 || p->cascade_order > 1)
    return fail(c, SWE_EINVAL, "cascade_order must be -1 (automatic), 0 (synchronous) or 1 (colour-ordered)");
  if (p->cascade_order == 0 && p->settling > swe::kSettlingMaxSync)
    return fail(c, SWE_EINVAL,
                "settling above 0.25 overshoots with the synchronous cascade: use cascade_order 1 (colour-ordered) "
                "or -1 (automatic)");
  return 0;
}

// the colour-ordered cascade in effect (cascade_order 1, or automatic above
// the synchronous cascade's settling bound)
bool colour_mode(const swe_params& p) {
  return p.cascade_order == 1 || (p.cascade_order < 0 && p.settling > swe::kSettlingMaxSync);
}

int ensure_layers(swe_ctx* c) {
  int r;
  if (c->p.layers == 2 && !c->smv[0]) {
    if ((r = dalloc(c, &c->smv[0], c->ncell)) || (r = dalloc(c, &c->smv[1], c->ncell)) ||
        (r = dalloc(c, &c->ss[0], c->ncell)) || (r = dalloc(c, &c->ss[1], c->ncell)) ||
        (r = dalloc(c, &c->ds, c->ncell)))
      return r;
  }
  if (c->p.track && !c->trk) {
    if ((r = dalloc(c, &c->trk, 4 * c->ncell))) return r;
  }
  return 0;
}

int ensure_particles(swe_ctx* c, int64_t n) {
  if (n <= c->cap) return 0;
  dfree(c->P0);
  dfree(c->P1);
  dfree(c->L);
  dfree(c->LC);
  c->cap = 0;
  int r;
  if ((r = dalloc(c, &c->P0, n))) return r;
  if ((r = dalloc(c, &c->P1, n))) return r;
  if ((r = dalloc(c, &c->L, n))) return r;
  if ((r = dalloc(c, &c->LC, n))) return r;
  c->cap = n;
  return 0;
}

// SWE_TUNE_HEAVY_FIRST in effect.  By default light tiles also go last when
// a launch is a few rounds of workgroups (up to 4096 tiles: 3.2 rounds at
// five 16x32 tiles per CU), where the order of the last round sets the
// launch's tail; with many rounds the tail matters little and moving tiles
// costs locality (measured: 1024^2 map +2.4%, 64-terrain batch -4%).
int32_t heavy_first_mode(const swe_ctx* c) {
  return c->heavy_first >= 0 ? c->heavy_first : (c->ntiles <= 4096 ? 2 : 1);
}

// light-tile slots of a launch: half the tiles, a multiple of 8
int64_t light_cap(int64_t ntiles) { return (ntiles / 2 + 7) / 8 * 8; }

// Tiles of every terrain for step-kernel shape `shape`: per-terrain tile
// grids, the per-tile terrain index on the device, and the per-tile routing
// buffers of the previous shape released (ensure_step makes new ones).
int set_tiles(swe_ctx* c, int32_t shape) {
  const swe::StepShape sh = swe::kStepShape[shape];
  int64_t tiles = 0;
  for (int32_t k = 0; k < c->nterr; k++)
    tiles += (int64_t)((c->tnx[(size_t)k] + sh.sx - 1) / sh.sx) * ((c->tny[(size_t)k] + sh.sy - 1) / sh.sy);
  if (tiles > ((int64_t)1 << 30)) return fail(c, SWE_EINVAL, "too many tiles");
  std::vector<int32_t> tile_terrain((size_t)tiles);
  std::vector<int32_t> ttx, tty, ttile0;
  int32_t tile0 = 0;
  for (int32_t k = 0; k < c->nterr; k++) {
    const int32_t tx = (c->tnx[(size_t)k] + sh.sx - 1) / sh.sx, ty = (c->tny[(size_t)k] + sh.sy - 1) / sh.sy;
    ttx.push_back(tx);
    tty.push_back(ty);
    ttile0.push_back(tile0);
    for (int32_t t = 0; t < tx * ty; t++) tile_terrain[(size_t)(tile0 + t)] = k;
    tile0 += tx * ty;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));  // no launch may still use the old buffers
  dfree(c->dtile_terrain);
  for (int k = 0; k < 2; k++) {
    dfree(c->sbase[k]);
    dfree(c->scnt[k]);
    dfree(c->heavy[k]);
  }
  dfree(c->nin);
  dfree(c->csum);
  dfree(c->hcnt);
  dfree(c->hstamp);
  int r;
  if ((r = dalloc(c, &c->dtile_terrain, tiles))) return r;
  HIPCHK(c, hipMemcpyAsync(c->dtile_terrain, tile_terrain.data(), sizeof(int32_t) * tile_terrain.size(),
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->ttx = ttx;
  c->tty = tty;
  c->ttile0 = ttile0;
  c->ntiles = (int32_t)tiles;
  c->shape = shape;
  return 0;
}

// buffers of the tile-owned step path for n particles
int ensure_step(swe_ctx* c, int64_t n) {
  int r;
  // particle indices and record positions are 32-bit on this path
  if (n > (int64_t)UINT32_MAX) return fail(c, SWE_EINVAL, "at most 2^32 - 1 particles per cycle on the tile step path");
  if (!c->fV[0]) {
    if ((r = dalloc(c, &c->fV[0], c->ncell)) || (r = dalloc(c, &c->fV[1], c->ncell)) ||
        (r = dalloc(c, &c->farcnt, 3)))
      return r;
  }
  if (!c->sbase[0]) {  // per-tile buffers of the current tile shape
    const int64_t nt = c->ntiles, nch = (nt + swe::kStepChunk - 1) / swe::kStepChunk;
    if ((r = dalloc(c, &c->sbase[0], 9 * nt)) ||
        (r = dalloc(c, &c->sbase[1], 9 * nt)) || (r = dalloc(c, &c->scnt[0], 9 * nt)) ||
        (r = dalloc(c, &c->scnt[1], 9 * nt)) || (r = dalloc(c, &c->nin, 3 * nt)) ||
        (r = dalloc(c, &c->csum, 3 * nch)) || (r = dalloc(c, &c->heavy[0], swe::kHeavySlotsMax + light_cap(nt))) ||
        (r = dalloc(c, &c->heavy[1], swe::kHeavySlotsMax + light_cap(nt))) || (r = dalloc(c, &c->hcnt, 48)) ||
        (r = dalloc(c, &c->hstamp, nt)))
      return r;
  }
  if (n > c->fcap) {
    for (int k = 0; k < 2; k++) dfree(c->PB[k]);
    for (int k = 0; k < 3; k++) dfree(c->FB[k]);
    dfree(c->sp_tile);
    dfree(c->sp_rank);
    c->fcap = 0;
    // a tile's output region has room for all its particles in each of its
    // nine direction sub-ranges
    for (int k = 0; k < 2; k++)
      if ((r = dalloc(c, &c->PB[k], 9 * n))) return r;
    for (int k = 0; k < 3; k++)
      if ((r = dalloc(c, &c->FB[k], n))) return r;
    if ((r = dalloc(c, &c->sp_tile, n)) || (r = dalloc(c, &c->sp_rank, n))) return r;
    c->fcap = n;
  }
  return 0;
}

hipEvent_t take_event(swe_ctx* c) {
  if (c->ev_used == c->ev_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    c->ev_pool.push_back(e);
  }
  return c->ev_pool[c->ev_used++];
}

// fold finished timing pairs into the stats (synchronizes the stream)
int collect_timing(swe_ctx* c) {
  if (c->pending.empty()) return 0;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (auto& pd : c->pending) {
    float ms = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&ms, pd.a, pd.b));
    if (pd.kind == 0 || pd.kind == 2) {
      c->st.march_ms += ms;
      c->st.march_launches++;
    } else {
      c->st.relax_ms += ms;
      c->st.relax_launches++;
    }
  }
  c->pending.clear();
  c->ev_used = 0;
  return 0;
}

int64_t windows_of(const swe_params& p, int32_t* R) {
  int64_t calls = (int64_t)p.max_age + 2;  // 513 particle-steps + the dying call
  if (calls < 1) calls = 1;
  const int32_t r = p.window > 0 ? p.window : (int32_t)calls;
  *R = r;
  return (calls + r - 1) / r;
}

void fill_consts(const swe_ctx* c, float ps[3], float nps[3]) {
  for (int k = 0; k < 3; k++) ps[k] = c->p.pspeed[k];
  normalize3(c->p.pspeed, nps);
}

int timed_begin(swe_ctx* c, hipEvent_t* e0, hipEvent_t* e1) {
  if (!c->timing) return 0;
  if ((c->timing_tick++ % (uint64_t)c->timing) != 0) return 0;  // not sampled: e0 stays null
  *e0 = take_event(c);
  *e1 = take_event(c);
  if (!*e0 || !*e1) return fail(c, SWE_EHIP, "hipEventCreate failed");
  HIPCHK(c, hipEventRecord(*e0, c->stream));
  return 0;
}

int timed_end(swe_ctx* c, hipEvent_t e0, hipEvent_t e1, int kind) {
  if (!c->timing || !e0) return 0;
  HIPCHK(c, hipEventRecord(e1, c->stream));
  c->pending.push_back({e0, e1, kind});
  return 0;
}

// World::erode (world.hpp:36-77) with the global-memory march: k_march
// (R fly() calls per particle, device-scope atomics) then k_relax per window
int erode_global(swe_ctx* c, int64_t particles, int32_t n_cycles, int64_t nwin, int32_t R) {
  int r;
  if ((r = ensure_particles(c, particles))) return r;
  swe::MarchArgs ma{};
  ma.n = particles;
  ma.ncell = c->ncell;
  ma.nx = c->nx;
  ma.ny = c->ny;
  ma.rcp_nx = 1.0f / (float)c->nx;
  ma.rcp_ny = 1.0f / (float)c->ny;
  ma.mapscale = c->p.mapscale;
  fill_consts(c, ma.ps, ma.nps);
  ma.p = c->p;
  ma.P0 = c->P0;
  ma.P1 = c->P1;
  ma.L = c->L;
  ma.LC = c->LC;
  ma.lost = c->lost;
  ma.stat = c->dstat;
  ma.cycle = c->dcycle;
  ma.T = c->trk;
  ma.D = c->dz;
  ma.DS = c->ds;
  ma.R = R;

  swe::RelaxArgs ra{};
  ra.nx = c->nx;
  ra.ny = c->ny;
  ra.rcp_nx = ma.rcp_nx;
  ra.rcp_ny = ma.rcp_ny;
  ra.maxdiff = c->p.maxdiff;
  ra.settling = c->p.settling;
  ra.dz = c->dz;
  ra.ds = c->ds;
  const bool layers2 = c->p.layers == 2;

  for (int32_t cyc = 0; cyc < n_cycles; cyc++) {
    if (c->p.track) swe::launch_zero64(c->trk, 4 * c->ncell, c->stream);
    for (int64_t w = 0; w < nwin; w++) {
      const int cur = c->cur();
      ma.S = c->zs[cur];
      ma.SS = layers2 ? c->ss[cur] : nullptr;
      ma.V = c->vis[cur];
      hipEvent_t e0 = nullptr, e1 = nullptr;
      if ((r = timed_begin(c, &e0, &e1))) return r;
      swe::launch_march(ma, w == 0, c->stream);
      HIPCHK(c, hipGetLastError());
      if ((r = timed_end(c, e0, e1, 0))) return r;
      ra.S = c->zs[cur];
      ra.SS = layers2 ? c->ss[cur] : nullptr;
      ra.V = c->vis[cur];
      ra.v_clear = c->vis[cur ^ 1];
      ra.zm_in = c->zmv[cur];
      ra.zm_out = c->zmv[cur ^ 1];
      ra.zs_out = c->zs[cur ^ 1];
      ra.sm_in = layers2 ? c->smv[cur] : nullptr;
      ra.sm_out = layers2 ? c->smv[cur ^ 1] : nullptr;
      ra.ss_out = layers2 ? c->ss[cur ^ 1] : nullptr;
      if ((r = timed_begin(c, &e0, &e1))) return r;
      if (colour_mode(c->p)) swe::launch_relax_colour(ra, c->p.periodic != 0, c->p.layers, true, c->stream);
      else swe::launch_relax(ra, c->p.periodic != 0, c->p.layers, c->stream);
      HIPCHK(c, hipGetLastError());
      if ((r = timed_end(c, e0, e1, 1))) return r;
      c->g++;
    }
    if (c->p.track) swe::launch_ema(c->ema, c->trk, 4 * c->ncell, c->p.lrate, c->stream);
    swe::launch_bump(c->dcycle, c->stream);
    HIPCHK(c, hipGetLastError());
    c->cycle++;
    c->st.cycles++;
    if (c->timing && c->pending.size() > 512) {
      if ((r = collect_timing(c))) return r;
    }
  }
  return 0;
}

// terrain records for the device: seed / pspeed from the params unless a
// terrain overrides them
int push_terrains(swe_ctx* c, const int64_t* particles) {
  std::vector<swe::Terrain> t((size_t)c->nterr);
  int64_t part0 = 0;
  for (int32_t k = 0; k < c->nterr; k++) {
    swe::Terrain& T = t[(size_t)k];
    T.cell0 = c->tcell0[(size_t)k];
    T.part0 = part0;
    T.npart = particles[k];
    part0 += particles[k];
    T.nx = c->tnx[(size_t)k];
    T.ny = c->tny[(size_t)k];
    T.tiles_x = c->ttx[(size_t)k];
    T.tiles_y = c->tty[(size_t)k];
    T.tile0 = c->ttile0[(size_t)k];
    T.rcp_nx = 1.0f / (float)T.nx;
    T.rcp_ny = 1.0f / (float)T.ny;
    T.seed = c->tseed_set[(size_t)k] ? c->tseed[(size_t)k] : c->p.seed;
    const float* ps = c->tps_set[(size_t)k] ? c->tps[(size_t)k].data() : c->p.pspeed;
    for (int q = 0; q < 3; q++) T.ps[q] = ps[q];
    normalize3(ps, T.nps);
  }
  HIPCHK(c, hipMemcpyAsync(c->dterr, t.data(), sizeof(swe::Terrain) * t.size(), hipMemcpyHostToDevice, c->stream));
  return 0;
}

// World::erode with one-call windows on the tile-owned step kernel, for every
// terrain of the context at once: spawn and bin by tile, then one k_step per
// window (relax of the previous window + one fly() call of every particle),
// then a relax-only k_step
int erode_tiled(swe_ctx* c, const int64_t* particles, int32_t n_cycles, int64_t nwin) {
  int r;
  int64_t total = 0;
  for (int32_t k = 0; k < c->nterr; k++) total += particles[k];
  if ((r = ensure_step(c, total > 0 ? total : 1))) return r;
  if ((r = push_terrains(c, particles))) return r;
  const bool layers2 = c->p.layers == 2;
  swe::StepArgs sa{};
  sa.trk = c->trk;
  sa.lost = c->lost;
  sa.stat = c->dstat;
  sa.cycle = c->dcycle;
  sa.terr = c->dterr;
  sa.tile_terrain = c->dtile_terrain;
  sa.nterr = c->nterr;
  sa.ncell = c->ncell;
  sa.ntiles = c->ntiles;
  sa.mapscale = c->p.mapscale;
  sa.boundarylayer = c->p.boundarylayer;
  sa.suspension = c->p.suspension;
  sa.gravity = c->p.gravity;
  sa.abrasion = c->p.abrasion;
  sa.maxdiff = c->p.maxdiff;
  sa.settling = c->p.settling;
  sa.max_age = c->p.max_age;
  sa.tile_order = c->tile_order;
  // listed crowded tiles get the first workgroups of a launch; at most one
  // slot per 8 tiles, a multiple of 8 (none on maps under 64 tiles)
  const int32_t hf = heavy_first_mode(c);
  sa.heavy_slots = hf ? (int32_t)std::min<int64_t>(swe::kHeavySlotsMax, (c->ntiles / 8) & ~7LL) : 0;
  // light tiles last (heavy-first mode 2, maps of 64 tiles or more): the
  // usual order is padded to a multiple of 8 so that slot b runs on XCD b % 8
  const bool light = hf == 2 && sa.heavy_slots > 0;
  sa.light_slots = light ? (int32_t)light_cap(c->ntiles) : 0;
  sa.default_slots = light ? (int32_t)((c->ntiles + 7) / 8 * 8) : c->ntiles;

  const int64_t nt = c->ntiles, nch = (nt + swe::kStepChunk - 1) / swe::kStepChunk;
  swe::SpawnArgs sp{};
  sp.out = c->PB[0];
  sp.tile = c->sp_tile;
  sp.rank = c->sp_rank;
  sp.count = c->nin;  // window 0's input counts
  sp.csum = c->csum;
  sp.sub_base = c->sbase[1];
  sp.sub_cnt = c->scnt[1];
  sp.cycle = c->dcycle;
  sp.terr = c->dterr;
  sp.nterr = c->nterr;
  sp.n = total;
  sp.ntiles = c->ntiles;
  sp.periodic = c->p.periodic;
  sp.spawn_mode = c->p.spawn;
  sp.tile_sx = swe::kStepShape[c->shape].sx;
  sp.tile_sy = swe::kStepShape[c->shape].sy;

  // colour-ordered cascade: every window's step kernel only moves the
  // particles (MODE 2), and its relax follows at once as one launch per
  // colour class; no closing relax-only window
  const bool colour = colour_mode(c->p);
  swe::RelaxArgs ra{};
  ra.nx = c->nx;
  ra.ny = c->ny;
  ra.rcp_nx = 1.0f / (float)c->nx;
  ra.rcp_ny = 1.0f / (float)c->ny;
  ra.maxdiff = c->p.maxdiff;
  ra.settling = c->p.settling;

  // the launches of one map-cycle (kernels only: no copies, fills or host
  // synchronisation, so that the sequence can be captured into a graph)
  auto enqueue_cycle = [&]() -> int {
    if (c->p.track) swe::launch_zero64(c->trk, 4 * c->ncell, c->stream);
    if (total > 0) {
      swe::launch_zero32(c->nin, 3 * nt, c->stream);
      swe::launch_zero32(c->csum, 3 * nch, c->stream);
      swe::launch_zero32(c->farcnt, 3, c->stream);
      if (sa.heavy_slots > 0) {
        swe::launch_zero32(c->hcnt, 48, c->stream);
        swe::launch_zero32(c->hstamp, nt, c->stream);
      }
      swe::launch_spawn(sp, c->stream);
      HIPCHK(c, hipGetLastError());
      // windows 0 .. nwin-1 step the particles; window nwin only relaxes
      for (int64_t w = 0; w <= nwin - (colour ? 1 : 0); w++) {
        const int cur = c->cur(), nxt = cur ^ 1;
        const int wp = (int)(w & 1);
        sa.zm_in = c->zmv[cur];
        sa.zm_out = c->zmv[nxt];
        sa.snap_in = c->zs[cur];
        sa.snap_out = c->zs[nxt];
        sa.sm_in = layers2 ? c->smv[cur] : nullptr;
        sa.sm_out = layers2 ? c->smv[nxt] : nullptr;
        sa.ssnap_in = layers2 ? c->ss[cur] : nullptr;
        sa.ssnap_out = layers2 ? c->ss[nxt] : nullptr;
        sa.v_in = c->fV[cur];
        sa.v_out = c->fV[nxt];
        const int w3 = (int)(w % 3), w3n = (int)((w + 1) % 3), w3c = (int)((w + 2) % 3);
        sa.in = c->PB[wp];
        sa.out = c->PB[wp ^ 1];
        sa.far_in = c->FB[w3];
        sa.far_out = c->FB[w3n];
        sa.far_cnt_in = c->farcnt + w3;
        sa.far_cnt_out = c->farcnt + w3n;
        sa.far_cnt_clear = c->farcnt + w3c;
        sa.sub_base_prev = c->sbase[wp ^ 1];
        sa.sub_cnt_prev = c->scnt[wp ^ 1];
        sa.sub_base_cur = c->sbase[wp];
        sa.sub_cnt_cur = c->scnt[wp];
        sa.nin_cur = c->nin + w3 * nt;
        sa.nin_next = c->nin + w3n * nt;
        sa.nin_clear = c->nin + w3c * nt;
        sa.csum_cur = c->csum + w3 * nch;
        sa.csum_next = c->csum + w3n * nch;
        sa.csum_clear = c->csum + w3c * nch;
        sa.window = (int32_t)w;
        sa.heavy_cur = c->heavy[wp];
        sa.heavy_next = c->heavy[wp ^ 1];
        sa.hcnt_cur = c->hcnt + 16 * w3;
        sa.hcnt_next = c->hcnt + 16 * w3n;
        sa.hcnt_clear = c->hcnt + 16 * w3c;
        sa.hstamp = c->hstamp;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        int rr;
        if ((rr = timed_begin(c, &e0, &e1))) return rr;
        swe::launch_step(c->shape, sa, c->p.periodic != 0, c->p.track != 0, c->p.layers,
                         c->p.deposit_on_death != 0, colour ? 2 : (w == nwin ? 1 : 0), c->stream);
        HIPCHK(c, hipGetLastError());
        if ((rr = timed_end(c, e0, e1, 2))) return rr;
        if (colour) {  // the step kernel wrote the masters with the pickups and their snapshots
          ra.V = c->fV[nxt];
          ra.zm_out = c->zmv[nxt];
          ra.zs_out = c->zs[nxt];
          ra.sm_out = layers2 ? c->smv[nxt] : nullptr;
          ra.ss_out = layers2 ? c->ss[nxt] : nullptr;
          if ((rr = timed_begin(c, &e0, &e1))) return rr;
          if (c->nterr > 1) {
            const int64_t per = (int64_t)swe::kColourAxisMax * swe::kColourAxisMax * (c->nterr + 1);
            const int64_t off = c->p.periodic ? 0 : per;
            swe::launch_relax_colour_batch(ra, c->dterr, c->dcolour + off, c->hcolour.data() + off, c->nterr,
                                           c->ncell, c->p.periodic != 0, c->p.layers, c->stream);
          } else {
            swe::launch_relax_colour(ra, c->p.periodic != 0, c->p.layers, false, c->stream);
          }
          // a cycle ends with no events pending, as after a relax-only window
          if (w == nwin - 1) swe::launch_zero32(c->fV[nxt], c->ncell, c->stream);
          HIPCHK(c, hipGetLastError());
          if ((rr = timed_end(c, e0, e1, 1))) return rr;
        }
        c->g++;
      }
    }
    if (c->p.track) swe::launch_ema(c->ema, c->trk, 4 * c->ncell, c->p.lrate, c->stream);
    swe::launch_bump(c->dcycle, c->stream);
    HIPCHK(c, hipGetLastError());
    return 0;
  };

  // graph mode: the cycle is captured once per starting parity and launched
  // as one graph; the capture is redone when anything it depends on changes
  const bool graph = c->use_graph != 0 && c->timing == 0;
  std::string key;
  if (graph) {
    key.append((const char*)&c->p, sizeof(c->p));
    key.append((const char*)particles, sizeof(int64_t) * (size_t)c->nterr);
    const void* ptrs[2] = {(const void*)c->PB[0], (const void*)c->trk};
    key.append((const char*)ptrs, sizeof(ptrs));
    key.append((const char*)&c->tile_order, sizeof(c->tile_order));
    key.append((const char*)&c->heavy_first, sizeof(c->heavy_first));
    key.append((const char*)&c->shape, sizeof(c->shape));
    key.append((const char*)&nwin, sizeof(nwin));
    key.append((const char*)&c->alloc_gen, sizeof(c->alloc_gen));
  }
  const uint64_t windows_per_cycle = total > 0 ? (uint64_t)nwin + (colour ? 0 : 1) : 0;

  for (int32_t cyc = 0; cyc < n_cycles; cyc++) {
    if (graph) {
      const int par = (int)(c->g & 1);
      if (!c->gexec[par] || c->gkey[par] != key) {
        if (c->gexec[par]) (void)hipGraphExecDestroy(c->gexec[par]);
        c->gexec[par] = nullptr;
        const uint64_t g0 = c->g;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_cycle();
        hipGraph_t gr = nullptr;
        const hipError_t ee = hipStreamEndCapture(c->stream, &gr);
        c->g = g0;  // the capture only recorded the launches
        if (rc != 0 || ee != hipSuccess) {
          if (gr) (void)hipGraphDestroy(gr);
          return rc != 0 ? rc : fail(c, SWE_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ee));
        }
        const hipError_t ei = hipGraphInstantiate(&c->gexec[par], gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (ei != hipSuccess) {
          c->gexec[par] = nullptr;
          return fail(c, SWE_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        }
        c->gkey[par] = key;
      }
      HIPCHK(c, hipGraphLaunch(c->gexec[par], c->stream));
      c->g += windows_per_cycle;
    } else {
      if ((r = enqueue_cycle())) return r;
    }
    c->cycle++;
    c->st.cycles++;
    if (c->timing && c->pending.size() > 512) {
      if ((r = collect_timing(c))) return r;
    }
  }
  return 0;
}

}  // namespace

extern "C" {

void swe_params_default(swe_params* p) {
  std::memset(p, 0, sizeof(*p));
  p->max_age = 512;
  p->boundarylayer = 2.0f;
  p->suspension = 0.05f;
  p->gravity = 0.1f;
  p->pspeed[0] = 1.0f;  // glm::normalize(vec3(1,0,0)) == (1,0,0) exactly
  p->pspeed[1] = 0.0f;
  p->pspeed[2] = 0.0f;
  p->lrate = 0.1f;
  p->maxdiff = 0.005f;
  p->settling = 1.0f;
  p->abrasion = 0.0f;
  p->mapscale = 80.0f;
  p->window = 1;
  p->periodic = 0;
  p->track = 1;
  p->layers = 1;
  p->deposit_on_death = 0;
  p->spawn = 0;
  p->seed = 0;
  p->cascade_order = -1;
  p->reserved = 0;
}

const char* swe_version(void) { return "swe-mi355x 0.1 (gfx950)"; }

const char* swe_last_error(const swe_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int swe_destroy(swe_ctx* c) {
  if (!c) return SWE_EINVAL;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (int k = 0; k < 2; k++) {
    dfree(c->zmv[k]);
    dfree(c->zs[k]);
    dfree(c->smv[k]);
    dfree(c->ss[k]);
    dfree(c->vis[k]);
    dfree(c->fV[k]);
    dfree(c->sbase[k]);
    dfree(c->scnt[k]);
    dfree(c->PB[k]);
  }
  for (int k = 0; k < 3; k++) dfree(c->FB[k]);
  dfree(c->trk);
  dfree(c->ema);
  dfree(c->dz);
  dfree(c->ds);
  dfree(c->P0);
  dfree(c->P1);
  dfree(c->L);
  dfree(c->LC);
  dfree(c->farcnt);
  dfree(c->nin);
  dfree(c->csum);
  dfree(c->heavy[0]);
  dfree(c->heavy[1]);
  dfree(c->hcnt);
  dfree(c->hstamp);
  dfree(c->sp_tile);
  dfree(c->sp_rank);
  dfree(c->lost);
  dfree(c->dstat);
  dfree(c->dcycle);
  dfree(c->dterr);
  dfree(c->dcolour);
  dfree(c->dtile_terrain);
  for (int k = 0; k < 2; k++)
    if (c->gexec[k]) (void)hipGraphExecDestroy(c->gexec[k]);
  for (auto e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

int swe_create_batch(swe_ctx** out, int32_t n_terrains, const int32_t* nx, const int32_t* ny, int32_t device) {
  if (!out) return SWE_EINVAL;
  *out = nullptr;
  if (n_terrains < 1 || n_terrains > 65536 || !nx || !ny) return SWE_EINVAL;
  int64_t cells = 0;
  for (int32_t k = 0; k < n_terrains; k++) {
    if (nx[k] < 1 || ny[k] < 1) return SWE_EINVAL;
    if ((int64_t)nx[k] * ny[k] > ((int64_t)1 << 31) - 1) return SWE_EINVAL;  // cell index fits int32
    cells += (int64_t)nx[k] * ny[k];
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SWE_EHIP;
  if (device < 0 || device >= ndev) return SWE_EINVAL;
  swe_ctx* c = new swe_ctx();
  c->device = device;
  c->nterr = n_terrains;
  c->nx = nx[0];
  c->ny = ny[0];
  c->ncell = cells;
  int64_t cell0 = 0;
  for (int32_t k = 0; k < n_terrains; k++) {
    c->tnx.push_back(nx[k]);
    c->tny.push_back(ny[k]);
    c->tcell0.push_back(cell0);
    cell0 += (int64_t)nx[k] * ny[k];
  }
  c->tseed.assign((size_t)n_terrains, 0);
  c->tps.assign((size_t)n_terrains, {1.0f, 0.0f, 0.0f});
  c->tseed_set.assign((size_t)n_terrains, 0);
  c->tps_set.assign((size_t)n_terrains, 0);
  swe_params_default(&c->p);
  int r = 0;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      resolve_scratch(device) != hipSuccess) {
    swe_destroy(c);
    return SWE_EHIP;
  }
  if ((r = dalloc(c, &c->zmv[0], c->ncell)) || (r = dalloc(c, &c->zmv[1], c->ncell)) ||
      (r = dalloc(c, &c->zs[0], c->ncell)) || (r = dalloc(c, &c->zs[1], c->ncell)) ||
      (r = dalloc(c, &c->dz, c->ncell)) ||
      (r = dalloc(c, &c->vis[0], c->ncell)) || (r = dalloc(c, &c->vis[1], c->ncell)) ||
      (r = dalloc(c, &c->ema, 4 * c->ncell)) || (r = dalloc(c, &c->lost, (int64_t)swe::kLostSlots * n_terrains)) ||
      (r = dalloc(c, &c->dstat, (int64_t)swe::kStatKinds * swe::kStatSlots)) ||
      (r = dalloc(c, &c->dcycle, 1)) || (r = dalloc(c, &c->dterr, n_terrains)) || (r = set_tiles(c, 0)) ||
      (r = ensure_layers(c))) {
    swe_destroy(c);
    return r;
  }
  if (n_terrains > 1) {
    const int64_t per = (int64_t)swe::kColourAxisMax * swe::kColourAxisMax * (n_terrains + 1);
    c->hcolour.assign((size_t)(2 * per), 0);
    swe::colour_prefix(nx, ny, n_terrains, true, c->hcolour.data());
    swe::colour_prefix(nx, ny, n_terrains, false, c->hcolour.data() + per);
    if ((r = dalloc(c, &c->dcolour, 2 * per)) ||
        hipMemcpy(c->dcolour, c->hcolour.data(), sizeof(int64_t) * (size_t)(2 * per), hipMemcpyHostToDevice) !=
            hipSuccess) {
      swe_destroy(c);
      return r ? r : SWE_EHIP;
    }
  }
  *out = c;
  return 0;
}

int swe_create(swe_ctx** out, int32_t nx, int32_t ny, int32_t device) {
  return swe_create_batch(out, 1, &nx, &ny, device);
}

int swe_set_terrain(swe_ctx* c, int32_t k, uint64_t seed, const float* pspeed) {
  if (!c) return SWE_EINVAL;
  if (k < 0 || k >= c->nterr) return fail(c, SWE_EINVAL, "terrain index out of range");
  c->tseed[(size_t)k] = seed;
  c->tseed_set[(size_t)k] = 1;
  if (pspeed) {
    for (int q = 0; q < 3; q++) c->tps[(size_t)k][q] = pspeed[q];
    c->tps_set[(size_t)k] = 1;
  }
  return 0;
}

int swe_get_terrains(const swe_ctx* c, int32_t* n_terrains) {
  if (!c || !n_terrains) return SWE_EINVAL;
  *n_terrains = c->nterr;
  return 0;
}

int swe_get_terrain(const swe_ctx* c, int32_t k, int32_t* nx, int32_t* ny, uint64_t* seed, float* pspeed,
                    int32_t* overrides) {
  if (!c || k < 0 || k >= c->nterr) return SWE_EINVAL;
  if (nx) *nx = c->tnx[(size_t)k];
  if (ny) *ny = c->tny[(size_t)k];
  const bool ss = c->tseed_set[(size_t)k] != 0, ps = c->tps_set[(size_t)k] != 0;
  if (seed) *seed = ss ? c->tseed[(size_t)k] : c->p.seed;
  if (pspeed)
    for (int q = 0; q < 3; q++) pspeed[q] = ps ? c->tps[(size_t)k][q] : c->p.pspeed[q];
  if (overrides) *overrides = (ss ? 1 : 0) | (ps ? 2 : 0);
  return 0;
}

int swe_set_params(swe_ctx* c, const swe_params* p) {
  if (!c || !p) return SWE_EINVAL;
  int r = check_params(c, p);
  if (r) return r;
  HIPCHK(c, hipSetDevice(c->device));
  const bool add_layer = p->layers == 2 && c->p.layers != 2 && !c->smv[0];
  c->p = *p;
  if ((r = ensure_layers(c))) return r;
  if (add_layer) {
    // a new sediment layer starts empty; its snapshot is the zeroed buffer
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return 0;
}

int swe_get_params(const swe_ctx* c, swe_params* p) {
  if (!c || !p) return SWE_EINVAL;
  *p = c->p;
  return 0;
}

int swe_upload_terrain(swe_ctx* c, int32_t k, const float* height, const float* sediment) {
  if (!c || !height) return SWE_EINVAL;
  if (k < 0 || k >= c->nterr) return fail(c, SWE_EINVAL, "terrain index out of range");
  HIPCHK(c, hipSetDevice(c->device));
  const int64_t c0 = c->tcell0[(size_t)k], n = (int64_t)c->tnx[(size_t)k] * c->tny[(size_t)k];
  if (sediment && !c->smv[0]) {
    // a sediment layer can be uploaded before layers == 2 is selected
    int r;
    if ((r = dalloc(c, &c->smv[0], c->ncell)) || (r = dalloc(c, &c->smv[1], c->ncell)) ||
        (r = dalloc(c, &c->ss[0], c->ncell)) || (r = dalloc(c, &c->ss[1], c->ncell)) ||
        (r = dalloc(c, &c->ds, c->ncell)))
      return r;
  }
  float* tmp = nullptr;
  HIPCHK(c, hipMalloc((void**)&tmp, sizeof(float) * n));
  hipError_t e = hipMemcpyAsync(tmp, height, sizeof(float) * n, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    swe::launch_from_f32(tmp, c->zmv[c->cur()] + c0, c->zs[c->cur()] + c0, n, c->stream);
    e = hipGetLastError();
  }
  if (e == hipSuccess && sediment) {
    e = hipMemcpyAsync(tmp, sediment, sizeof(float) * n, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
      swe::launch_from_f32(tmp, c->smv[c->cur()] + c0, c->ss[c->cur()] + c0, n, c->stream);
      e = hipGetLastError();
    }
  }
  hipError_t e2 = hipStreamSynchronize(c->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return fail(c, SWE_EHIP, std::string("upload: ") + hipGetErrorString(e));
  if (e2 != hipSuccess) return fail(c, SWE_EHIP, std::string("upload sync: ") + hipGetErrorString(e2));
  return 0;
}

int swe_upload_layers(swe_ctx* c, const float* height, const float* sediment) {
  if (!c) return SWE_EINVAL;
  if (c->nterr != 1) return fail(c, SWE_EINVAL, "a batch context uploads with swe_upload_terrain");
  return swe_upload_terrain(c, 0, height, sediment);
}

int swe_upload_terrain_master(swe_ctx* c, int32_t k, const int64_t* height_fx, const int64_t* sediment_fx) {
  if (!c) return SWE_EINVAL;
  if (k < 0 || k >= c->nterr) return fail(c, SWE_EINVAL, "terrain index out of range");
  HIPCHK(c, hipSetDevice(c->device));
  const int64_t c0 = c->tcell0[(size_t)k], n = (int64_t)c->tnx[(size_t)k] * c->tny[(size_t)k];
  if (sediment_fx && !c->smv[0]) return fail(c, SWE_ESTATE, "no sediment layer: set layers = 2 first");
  if (height_fx) {
    HIPCHK(c, hipMemcpyAsync(c->zmv[c->cur()] + c0, height_fx, sizeof(int64_t) * n, hipMemcpyHostToDevice, c->stream));
    swe::launch_snap(c->zmv[c->cur()] + c0, c->zs[c->cur()] + c0, n, c->stream);
    HIPCHK(c, hipGetLastError());
  }
  if (sediment_fx) {
    HIPCHK(c, hipMemcpyAsync(c->smv[c->cur()] + c0, sediment_fx, sizeof(int64_t) * n, hipMemcpyHostToDevice, c->stream));
    swe::launch_snap(c->smv[c->cur()] + c0, c->ss[c->cur()] + c0, n, c->stream);
    HIPCHK(c, hipGetLastError());
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int swe_upload_master(swe_ctx* c, const int64_t* height_fx, const int64_t* sediment_fx) {
  if (!c) return SWE_EINVAL;
  if (c->nterr != 1) return fail(c, SWE_EINVAL, "a batch context uploads with swe_upload_terrain_master");
  return swe_upload_terrain_master(c, 0, height_fx, sediment_fx);
}

int swe_upload_terrain_fields(swe_ctx* c, int32_t k, const float* massflow, const float* momentum) {
  if (!c) return SWE_EINVAL;
  if (k < 0 || k >= c->nterr) return fail(c, SWE_EINVAL, "terrain index out of range");
  HIPCHK(c, hipSetDevice(c->device));
  const int64_t c0 = c->tcell0[(size_t)k], n = (int64_t)c->tnx[(size_t)k] * c->tny[(size_t)k];
  if (massflow) HIPCHK(c, hipMemcpyAsync(c->ema + c0, massflow, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
  if (momentum)
    for (int q = 0; q < 3; q++)
      HIPCHK(c, hipMemcpyAsync(c->ema + (q + 1) * c->ncell + c0, momentum + q * n, sizeof(float) * n,
                               hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int swe_upload_fields(swe_ctx* c, const float* massflow, const float* momentum) {
  if (!c) return SWE_EINVAL;
  if (c->nterr != 1) return fail(c, SWE_EINVAL, "a batch context uploads with swe_upload_terrain_fields");
  return swe_upload_terrain_fields(c, 0, massflow, momentum);
}

int swe_download_terrain(swe_ctx* c, int32_t k, float* height, float* sediment, float* massflow, float* momentum) {
  if (!c) return SWE_EINVAL;
  if (k < 0 || k >= c->nterr) return fail(c, SWE_EINVAL, "terrain index out of range");
  HIPCHK(c, hipSetDevice(c->device));
  const int64_t c0 = c->tcell0[(size_t)k], n = (int64_t)c->tnx[(size_t)k] * c->tny[(size_t)k];
  if (height) HIPCHK(c, hipMemcpyAsync(height, c->zs[c->cur()] + c0, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
  if (sediment) {
    if (c->ss[c->cur()])
      HIPCHK(c, hipMemcpyAsync(sediment, c->ss[c->cur()] + c0, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    else
      std::memset(sediment, 0, sizeof(float) * n);
  }
  if (massflow) HIPCHK(c, hipMemcpyAsync(massflow, c->ema + c0, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
  if (momentum)
    for (int q = 0; q < 3; q++)
      HIPCHK(c, hipMemcpyAsync(momentum + q * n, c->ema + (q + 1) * c->ncell + c0, sizeof(float) * n,
                               hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int swe_download_layers(swe_ctx* c, float* height, float* sediment, float* massflow, float* momentum) {
  if (!c) return SWE_EINVAL;
  if (c->nterr != 1) return fail(c, SWE_EINVAL, "a batch context downloads with swe_download_terrain");
  return swe_download_terrain(c, 0, height, sediment, massflow, momentum);
}

int swe_download_terrain_master(swe_ctx* c, int32_t k, int64_t* height_fx, int64_t* sediment_fx) {
  if (!c) return SWE_EINVAL;
  if (k < 0 || k >= c->nterr) return fail(c, SWE_EINVAL, "terrain index out of range");
  HIPCHK(c, hipSetDevice(c->device));
  const int64_t c0 = c->tcell0[(size_t)k], n = (int64_t)c->tnx[(size_t)k] * c->tny[(size_t)k];
  if (height_fx)
    HIPCHK(c, hipMemcpyAsync(height_fx, c->zmv[c->cur()] + c0, sizeof(int64_t) * n, hipMemcpyDeviceToHost, c->stream));
  if (sediment_fx) {
    if (c->smv[0])
      HIPCHK(c, hipMemcpyAsync(sediment_fx, c->smv[c->cur()] + c0, sizeof(int64_t) * n, hipMemcpyDeviceToHost,
                               c->stream));
    else
      std::memset(sediment_fx, 0, sizeof(int64_t) * n);
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int swe_download_master(swe_ctx* c, int64_t* height_fx, int64_t* sediment_fx) {
  if (!c) return SWE_EINVAL;
  if (c->nterr != 1) return fail(c, SWE_EINVAL, "a batch context downloads with swe_download_terrain_master");
  return swe_download_terrain_master(c, 0, height_fx, sediment_fx);
}

int swe_erode_batch(swe_ctx* c, const int64_t* particles, int32_t n_cycles) {
  if (!c || !particles) return SWE_EINVAL;
  if (n_cycles < 0) return fail(c, SWE_EINVAL, "n_cycles must be >= 0");
  int64_t total = 0;
  for (int32_t k = 0; k < c->nterr; k++) {
    if (particles[k] < 0) return fail(c, SWE_EINVAL, "particles must be >= 0");
    total += particles[k];
  }
  if (total > ((int64_t)1 << 31) - 256) return fail(c, SWE_EINVAL, "particles per cycle above 2^31");
  HIPCHK(c, hipSetDevice(c->device));
  int r;
  if ((r = ensure_layers(c))) return r;
  int32_t R;
  const int64_t nwin = windows_of(c->p, &R);
  const bool tiled = c->path == 2 || (c->path == 0 && R == 1) || c->nterr > 1;
  if (tiled && R != 1) return fail(c, SWE_EINVAL, "the tile-owned step path (and batches) need window == 1");
  if (tiled) return erode_tiled(c, particles, n_cycles, nwin);
  return erode_global(c, particles[0], n_cycles, nwin, R);
}

int swe_autotune(swe_ctx* c, const int64_t* particles, int32_t trial_cycles, swe_tune_result* out) {
  if (!c || !particles || !out) return SWE_EINVAL;
  if (trial_cycles < 1) return fail(c, SWE_EINVAL, "trial_cycles must be >= 1");
  HIPCHK(c, hipSetDevice(c->device));
  int r;
  if ((r = ensure_layers(c))) return r;
  int32_t R;
  (void)windows_of(c->p, &R);
  if (!(c->path == 2 || (c->path == 0 && R == 1) || c->nterr > 1))
    return fail(c, SWE_EINVAL, "autotune covers the tile step path (window == 1)");
  if ((r = collect_timing(c))) return r;
  // the state a trial changes, kept on the device and put back afterwards
  const int cur0 = c->cur();
  const uint64_t g0 = c->g, cycle0 = c->cycle;
  const swe_stats_t st0 = c->st;
  const int64_t n = c->ncell;
  const int64_t nlost = (int64_t)swe::kLostSlots * c->nterr, nstat = (int64_t)swe::kStatKinds * swe::kStatSlots;
  int64_t *zm = nullptr, *sm = nullptr, *lost = nullptr;
  float *zs = nullptr, *ss = nullptr, *ema = nullptr;
  uint64_t *cyc = nullptr, *stat = nullptr;
  const bool l2 = c->smv[0] != nullptr;
  auto release = [&]() {
    dfree(zm); dfree(sm); dfree(lost); dfree(zs); dfree(ss); dfree(ema); dfree(cyc); dfree(stat);
  };
  if ((r = dalloc(c, &zm, n)) || (r = dalloc(c, &zs, n)) || (l2 && ((r = dalloc(c, &sm, n)) || (r = dalloc(c, &ss, n)))) ||
      (r = dalloc(c, &ema, 4 * n)) || (r = dalloc(c, &lost, nlost)) || (r = dalloc(c, &cyc, 1)) ||
      (r = dalloc(c, &stat, nstat))) {
    release();
    return r;
  }
  auto copy = [&](void* dst, const void* src, size_t bytes) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream);
  };
  auto save_or_restore = [&](bool save) -> hipError_t {
    hipError_t e = hipSuccess;
    auto cp = [&](void* state, void* keep, size_t bytes) {
      if (e == hipSuccess) e = save ? copy(keep, state, bytes) : copy(state, keep, bytes);
    };
    cp(c->zmv[cur0], zm, sizeof(int64_t) * n);
    cp(c->zs[cur0], zs, sizeof(float) * n);
    if (l2) {
      cp(c->smv[cur0], sm, sizeof(int64_t) * n);
      cp(c->ss[cur0], ss, sizeof(float) * n);
    }
    cp(c->ema, ema, sizeof(float) * 4 * n);
    cp(c->lost, lost, sizeof(int64_t) * nlost);
    cp(c->dcycle, cyc, sizeof(uint64_t));
    cp(c->dstat, stat, sizeof(uint64_t) * nstat);
    if (!save && e == hipSuccess && c->fV[cur0]) {
      swe::launch_zero32(c->fV[cur0], n, c->stream);  // no events pending at a cycle start
      e = hipGetLastError();
    }
    c->g = g0;
    c->cycle = cycle0;
    c->st = st0;
    return e;
  };
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t he = save_or_restore(true);
  if (he == hipSuccess) he = hipEventCreate(&e0);
  if (he == hipSuccess) he = hipEventCreate(&e1);
  if (he != hipSuccess) {
    release();
    if (e0) (void)hipEventDestroy(e0);
    return fail(c, SWE_EHIP, std::string("autotune set-up: ") + hipGetErrorString(he));
  }
  const int32_t order0 = c->tile_order, graph0 = c->use_graph, heavy_set0 = c->heavy_first, timing0 = c->timing;
  const int32_t shape0 = c->shape;
  c->timing = 0;
  static const int32_t kOrders[] = {0, 1, 4, 16};
  double best = 1e300;
  int32_t ncand = 0;
  // one trial: an untimed cycle (graph capture, caches), then the timed
  // ones, each from the saved state; returns the ms per cycle or < 0
  auto trial = [&]() -> double {
    if ((he = save_or_restore(false)) != hipSuccess) return -1.0;
    if ((r = swe_erode_batch(c, particles, 1))) return -1.0;
    if ((he = hipEventRecord(e0, c->stream)) != hipSuccess) return -1.0;
    if ((r = swe_erode_batch(c, particles, trial_cycles))) return -1.0;
    if ((he = hipEventRecord(e1, c->stream)) != hipSuccess) return -1.0;
    if ((he = hipEventSynchronize(e1)) != hipSuccess) return -1.0;
    float ms = 0.0f;
    if ((he = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) return -1.0;
    ncand++;
    return ms / trial_cycles;
  };
  // the tile shape first, with graphs, the first tile order and the
  // default heavy-first mode
  int32_t best_shape = shape0;
  for (int32_t sh = 0; sh < swe::kStepShapes && r == 0 && he == hipSuccess; sh++) {
    if ((r = swe_set_tune(c, SWE_TUNE_SHAPE, sh))) break;
    c->tile_order = 0;
    c->use_graph = 1;
    c->heavy_first = -1;
    const double per = trial();
    if (per < 0) break;
    if (per < best) {
      best = per;
      best_shape = sh;
    }
  }
  if (r == 0 && he == hipSuccess) r = swe_set_tune(c, SWE_TUNE_SHAPE, best_shape);
  const int32_t heavy0 = heavy_first_mode(c);
  int32_t best_order = 0, best_graph = 1, best_heavy = -1;
  // heavy-first only changes anything where a launch has heavy slots
  const int32_t heavy_lo = c->ntiles / 8 >= 8 ? 0 : heavy0, heavy_hi = c->ntiles / 8 >= 8 ? 2 : heavy0;
  for (int32_t hv = heavy_hi; hv >= heavy_lo && r == 0 && he == hipSuccess; hv--)
  for (int32_t graph = 0; graph <= 1 && r == 0; graph++) {
    for (int32_t order : kOrders) {
      c->tile_order = order;
      c->use_graph = graph;
      c->heavy_first = hv;
      const double per = trial();
      if (per < 0) break;
      if (per < best) {
        best = per;
        best_order = order;
        best_graph = graph;
        best_heavy = hv;
      }
    }
    if (he != hipSuccess) break;
  }
  const hipError_t hr = save_or_restore(false);
  if (he == hipSuccess) he = hr;
  if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  release();
  c->timing = timing0;
  const bool ok = r == 0 && he == hipSuccess;
  c->tile_order = ok ? best_order : order0;
  c->use_graph = ok ? best_graph : graph0;
  c->heavy_first = ok ? best_heavy : heavy_set0;
  if (!ok && c->shape != shape0) (void)swe_set_tune(c, SWE_TUNE_SHAPE, shape0);
  if (r) return r;
  if (he != hipSuccess) return fail(c, SWE_EHIP, std::string("autotune: ") + hipGetErrorString(he));
  out->tile_order = c->tile_order;
  out->graph = c->use_graph;
  out->heavy_first = heavy_first_mode(c);
  out->candidates = ncand;
  out->ms_per_cycle = best;
  out->shape = c->shape;
  return 0;
}

int swe_device_arch(swe_ctx* c, char* buf, int32_t len) {
  if (!c || !buf || len < 1) return SWE_EINVAL;
  hipDeviceProp_t prop;
  HIPCHK(c, hipGetDeviceProperties(&prop, c->device));
  std::snprintf(buf, (size_t)len, "%s", prop.gcnArchName);
  return 0;
}

int swe_erode(swe_ctx* c, int64_t particles, int32_t n_cycles) {
  if (!c) return SWE_EINVAL;
  if (c->nterr != 1) return fail(c, SWE_EINVAL, "a batch context erodes with swe_erode_batch");
  if (particles < 0) return fail(c, SWE_EINVAL, "particles and n_cycles must be >= 0");
  return swe_erode_batch(c, &particles, n_cycles);
}

int swe_set_tune(swe_ctx* c, int32_t knob, int64_t value) {
  if (!c) return SWE_EINVAL;
  switch (knob) {
    case SWE_TUNE_TILE_ORDER:
      if (value < 0 || value > 4096) return fail(c, SWE_EINVAL, "tile order must be in [0, 4096]");
      c->tile_order = (int32_t)value;
      return 0;
    case SWE_TUNE_GRAPH:
      if (value != 0 && value != 1) return fail(c, SWE_EINVAL, "graph mode must be 0 or 1");
      c->use_graph = (int32_t)value;
      return 0;
    case SWE_TUNE_HEAVY_FIRST:
      if (value < 0 || value > 2) return fail(c, SWE_EINVAL, "heavy-first must be 0, 1 or 2");
      c->heavy_first = (int32_t)value;
      return 0;
    case SWE_TUNE_SHAPE:
      if (value < 0 || value >= swe::kStepShapes) return fail(c, SWE_EINVAL, "unknown step-kernel shape");
      if ((int32_t)value == c->shape) return 0;
      HIPCHK(c, hipSetDevice(c->device));
      return set_tiles(c, (int32_t)value);
    default:
      return fail(c, SWE_EINVAL, "unknown tuning knob");
  }
}

int swe_get_tune(swe_ctx* c, int32_t knob, int64_t* value) {
  if (!c || !value) return SWE_EINVAL;
  switch (knob) {
    case SWE_TUNE_TILE_ORDER:
      *value = c->tile_order;
      return 0;
    case SWE_TUNE_GRAPH:
      *value = c->use_graph;
      return 0;
    case SWE_TUNE_HEAVY_FIRST:
      *value = heavy_first_mode(c);
      return 0;
    case SWE_TUNE_SHAPE:
      *value = c->shape;
      return 0;
    default:
      return fail(c, SWE_EINVAL, "unknown tuning knob");
  }
}

int swe_set_path(swe_ctx* c, int32_t path) {
  if (!c) return SWE_EINVAL;
  if (path < 0 || path > 2) return fail(c, SWE_EINVAL, "path must be 0 (auto), 1 (global march) or 2 (tile step)");
  c->path = path;
  return 0;
}

int swe_synchronize(swe_ctx* c) {
  if (!c) return SWE_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return collect_timing(c);
}

int swe_audit_terrain(swe_ctx* c, int32_t k, swe_audit_t* out) {
  if (!c || !out) return SWE_EINVAL;
  if (k < 0 || k >= c->nterr) return fail(c, SWE_EINVAL, "terrain index out of range");
  HIPCHK(c, hipSetDevice(c->device));
  const int64_t c0 = c->tcell0[(size_t)k], n = (int64_t)c->tnx[(size_t)k] * c->tny[(size_t)k];
  std::vector<int64_t> z((size_t)n), s;
  int64_t lost[swe::kLostSlots];
  HIPCHK(c, hipMemcpyAsync(z.data(), c->zmv[c->cur()] + c0, sizeof(int64_t) * n, hipMemcpyDeviceToHost, c->stream));
  if (c->smv[0]) {
    s.resize((size_t)n);
    HIPCHK(c, hipMemcpyAsync(s.data(), c->smv[c->cur()] + c0, sizeof(int64_t) * n, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(c, hipMemcpyAsync(lost, c->lost + (int64_t)k * swe::kLostSlots, sizeof(lost), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  __int128 hz = 0, hs = 0;
  for (int64_t v : z) hz += v;
  for (int64_t v : s) hs += v;
  int64_t l = 0;
  for (int q = 0; q < swe::kLostSlots; q++) l += lost[q];
  out->height_sum_lo = (int64_t)(uint64_t)hz;
  out->height_sum_hi = (int64_t)(hz >> 64);
  out->sediment_sum_lo = (int64_t)(uint64_t)hs;
  out->sediment_sum_hi = (int64_t)(hs >> 64);
  out->lost = l;
  out->cycle = c->cycle;
  return 0;
}

int swe_audit(swe_ctx* c, swe_audit_t* out) {
  if (!c) return SWE_EINVAL;
  if (c->nterr != 1) return fail(c, SWE_EINVAL, "a batch context audits with swe_audit_terrain");
  return swe_audit_terrain(c, 0, out);
}

int swe_get_size(const swe_ctx* c, int32_t* nx, int32_t* ny) {
  if (!c || !nx || !ny) return SWE_EINVAL;
  *nx = c->nx;
  *ny = c->ny;
  return 0;
}

int swe_get_cycle(const swe_ctx* c, uint64_t* cycle) {
  if (!c || !cycle) return SWE_EINVAL;
  *cycle = c->cycle;
  return 0;
}

int swe_set_cycle(swe_ctx* c, uint64_t cycle) {
  if (!c) return SWE_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(c->dcycle, &cycle, sizeof(cycle), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->cycle = cycle;
  return 0;
}

int swe_set_terrain_lost(swe_ctx* c, int32_t k, int64_t lost) {
  if (!c) return SWE_EINVAL;
  if (k < 0 || k >= c->nterr) return fail(c, SWE_EINVAL, "terrain index out of range");
  HIPCHK(c, hipSetDevice(c->device));
  int64_t slots[swe::kLostSlots] = {0};
  slots[0] = lost;
  HIPCHK(c, hipMemcpyAsync(c->lost + (int64_t)k * swe::kLostSlots, slots, sizeof(slots), hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int swe_set_lost(swe_ctx* c, int64_t lost) {
  if (!c) return SWE_EINVAL;
  if (c->nterr != 1) return fail(c, SWE_EINVAL, "a batch context sets it with swe_set_terrain_lost");
  return swe_set_terrain_lost(c, 0, lost);
}

int swe_set_timing(swe_ctx* c, int32_t on) {
  if (!c) return SWE_EINVAL;
  int r = collect_timing(c);
  if (r) return r;
  if (on < 0) return fail(c, SWE_EINVAL, "timing stride must be >= 0");
  c->timing = on;
  c->timing_tick = 0;
  return 0;
}

int swe_stats(swe_ctx* c, swe_stats_t* out) {
  if (!c || !out) return SWE_EINVAL;
  int r = collect_timing(c);
  if (r) return r;
  *out = c->st;
  uint64_t h[swe::kStatKinds * swe::kStatSlots];
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(h, c->dstat, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  out->particle_steps = out->max_tile_particles = out->crowded_tile_windows = out->far_particles = 0;
  out->heavy_tile_windows = out->light_tile_windows = 0;
  for (int q = 0; q < swe::kStatSlots; q++) {
    out->particle_steps += h[swe::kStatSteps * swe::kStatSlots + q];
    const uint64_t m = h[swe::kStatMaxTile * swe::kStatSlots + q];
    if (m > out->max_tile_particles) out->max_tile_particles = m;
    out->crowded_tile_windows += h[swe::kStatCrowded * swe::kStatSlots + q];
    out->heavy_tile_windows += h[swe::kStatHeavy * swe::kStatSlots + q];
    out->light_tile_windows += h[swe::kStatLight * swe::kStatSlots + q];
    out->far_particles += h[swe::kStatFar * swe::kStatSlots + q];
  }
  return 0;
}

int swe_reset_stats(swe_ctx* c) {
  if (!c) return SWE_EINVAL;
  int r = collect_timing(c);
  if (r) return r;
  std::memset(&c->st, 0, sizeof(c->st));
  HIPCHK(c, hipSetDevice(c->device));
  swe::launch_zero64((int64_t*)c->dstat, (int64_t)swe::kStatKinds * swe::kStatSlots, c->stream);
  HIPCHK(c, hipGetLastError());
  return 0;
}

int swe_mesh(swe_ctx* c, float* out) {
  if (!c || !out) return SWE_EINVAL;
  if (c->nterr != 1) return fail(c, SWE_EINVAL, "not available on a batch context");
  HIPCHK(c, hipSetDevice(c->device));
  float* tmp = nullptr;
  HIPCHK(c, hipMalloc((void**)&tmp, sizeof(float) * 12 * c->ncell));
  swe::launch_mesh(c->zs[c->cur()], tmp, c->nx, c->ny, c->p.mapscale, c->p.periodic != 0, c->stream);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, tmp, sizeof(float) * 12 * c->ncell, hipMemcpyDeviceToHost, c->stream);
  hipError_t e2 = hipStreamSynchronize(c->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return fail(c, SWE_EHIP, std::string("mesh: ") + hipGetErrorString(e));
  if (e2 != hipSuccess) return fail(c, SWE_EHIP, std::string("mesh sync: ") + hipGetErrorString(e2));
  return 0;
}

int swe_windmap_terrain(swe_ctx* c, int32_t k, float* out) {
  if (!c || !out) return SWE_EINVAL;
  if (k < 0 || k >= c->nterr) return fail(c, SWE_EINVAL, "terrain index out of range");
  HIPCHK(c, hipSetDevice(c->device));
  const int64_t c0 = c->tcell0[(size_t)k], n = (int64_t)c->tnx[(size_t)k] * c->tny[(size_t)k];
  float* tmp = nullptr;
  HIPCHK(c, hipMalloc((void**)&tmp, sizeof(float) * 4 * n));
  // the momentum planes of terrain k: x at ema[ncell + c0], y and z a plane stride further
  swe::launch_windmap(c->ema + c->ncell + c0, c->ncell, tmp, n, c->stream);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, tmp, sizeof(float) * 4 * n, hipMemcpyDeviceToHost, c->stream);
  hipError_t e2 = hipStreamSynchronize(c->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return fail(c, SWE_EHIP, std::string("windmap: ") + hipGetErrorString(e));
  if (e2 != hipSuccess) return fail(c, SWE_EHIP, std::string("windmap sync: ") + hipGetErrorString(e2));
  return 0;
}

int swe_windmap(swe_ctx* c, float* out) {
  if (!c || !out) return SWE_EINVAL;
  if (c->nterr != 1) return fail(c, SWE_EINVAL, "a batch context uses swe_windmap_terrain");
  return swe_windmap_terrain(c, 0, out);
}

int swe_normals_terrain(swe_ctx* c, int32_t k, float* out) {
  if (!c || !out) return SWE_EINVAL;
  if (k < 0 || k >= c->nterr) return fail(c, SWE_EINVAL, "terrain index out of range");
  HIPCHK(c, hipSetDevice(c->device));
  const int32_t nx = c->tnx[(size_t)k], ny = c->tny[(size_t)k];
  const int64_t c0 = c->tcell0[(size_t)k], n = (int64_t)nx * ny;
  float* tmp = nullptr;
  HIPCHK(c, hipMalloc((void**)&tmp, sizeof(float) * 3 * n));
  swe::launch_normals(c->zs[c->cur()] + c0, tmp, nx, ny, c->p.mapscale, c->p.periodic != 0, c->stream);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, tmp, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, c->stream);
  hipError_t e2 = hipStreamSynchronize(c->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return fail(c, SWE_EHIP, std::string("normals: ") + hipGetErrorString(e));
  if (e2 != hipSuccess) return fail(c, SWE_EHIP, std::string("normals sync: ") + hipGetErrorString(e2));
  return 0;
}

int swe_normals(swe_ctx* c, float* out) {
  if (!c || !out) return SWE_EINVAL;
  if (c->nterr != 1) return fail(c, SWE_EINVAL, "a batch context uses swe_normals_terrain");
  return swe_normals_terrain(c, 0, out);
}

// Device memory of a call that may return early (HIPCHK): freed on scope exit
struct DevScratch {
  void* p = nullptr;
  ~DevScratch() {
    if (p) (void)hipFree(p);
  }
};

int swe_debug_step(swe_ctx* c, int32_t n, float* state, int32_t* age, const uint64_t* key, int32_t* alive,
                   int64_t* pickup) {
  if (!c || n < 0 || (n > 0 && (!state || !age || !key || !alive || !pickup))) return SWE_EINVAL;
  if (c->nterr != 1) return fail(c, SWE_EINVAL, "swe_debug_step: a map context only");
  if (n == 0) return 0;
  HIPCHK(c, hipSetDevice(c->device));
  int r;
  if ((r = ensure_layers(c))) return r;
  if (!c->zs[c->cur()] || (c->p.layers == 2 && !c->ss[c->cur()]))
    return fail(c, SWE_ESTATE, "swe_debug_step: no snapshot");
  // one allocation: state (7 floats), age, alive (int32), key, pick (64-bit) per particle
  const size_t b_st = sizeof(float) * 7 * (size_t)n, b_i = sizeof(int32_t) * (size_t)n,
               b_q = sizeof(uint64_t) * (size_t)n;
  const size_t o_age = (b_st + 15) / 16 * 16, o_alive = o_age + (b_i + 15) / 16 * 16,
               o_key = o_alive + (b_i + 15) / 16 * 16, o_pick = o_key + b_q, total = o_pick + b_q;
  DevScratch buf;
  if (hipMalloc(&buf.p, total) != hipSuccess) {
    buf.p = nullptr;
    return fail(c, SWE_ENOMEM, "swe_debug_step: hipMalloc failed");
  }
  char* base = (char*)buf.p;
  swe::DebugFlyArgs a{};
  a.S = c->zs[c->cur()];
  a.SS = c->p.layers == 2 ? c->ss[c->cur()] : nullptr;
  a.st = (float*)base;
  a.age = (int32_t*)(base + o_age);
  a.alive = (int32_t*)(base + o_alive);
  a.key = (const uint64_t*)(base + o_key);
  a.pick = (int64_t*)(base + o_pick);
  a.n = n;
  a.nx = c->nx;
  a.ny = c->ny;
  a.rcp_nx = 1.0f / (float)c->nx;
  a.rcp_ny = 1.0f / (float)c->ny;
  a.mapscale = c->p.mapscale;
  fill_consts(c, a.ps, a.nps);
  a.p = c->p;
  HIPCHK(c, hipMemcpyAsync(a.st, state, b_st, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(a.age, age, b_i, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync((void*)a.key, key, b_q, hipMemcpyHostToDevice, c->stream));
  swe::launch_debug_fly(a, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(state, a.st, b_st, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(age, a.age, b_i, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(alive, a.alive, b_i, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(pickup, a.pick, b_q, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int swe_debug_cascade(swe_ctx* c, int32_t x, int32_t y, uint32_t events, int32_t* count, int64_t* nb_cell,
                      int64_t* nb_fx) {
  if (!c || !count || !nb_cell || !nb_fx) return SWE_EINVAL;
  if (c->nterr != 1) return fail(c, SWE_EINVAL, "swe_debug_cascade: a map context only");
  if (x < 0 || y < 0 || x >= c->nx || y >= c->ny) return fail(c, SWE_EINVAL, "swe_debug_cascade: cell out of the map");
  HIPCHK(c, hipSetDevice(c->device));
  int r;
  if ((r = ensure_layers(c))) return r;
  if (!c->zs[c->cur()] || (c->p.layers == 2 && !c->ss[c->cur()]))
    return fail(c, SWE_ESTATE, "swe_debug_cascade: no snapshot");
  DevScratch buf;
  if (hipMalloc(&buf.p, sizeof(int64_t) * 17) != hipSuccess) {
    buf.p = nullptr;
    return fail(c, SWE_ENOMEM, "swe_debug_cascade: hipMalloc failed");
  }
  swe::DebugCascadeArgs a{};
  a.S = c->zs[c->cur()];
  a.SS = c->p.layers == 2 ? c->ss[c->cur()] : nullptr;
  a.nx = c->nx;
  a.ny = c->ny;
  a.x = x;
  a.y = y;
  a.events = events;
  a.cell = (int64_t*)buf.p;
  a.fx = a.cell + 8;
  a.count = (int32_t*)(a.cell + 16);
  a.p = c->p;
  swe::launch_debug_cascade(a, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(nb_cell, a.cell, sizeof(int64_t) * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(nb_fx, a.fx, sizeof(int64_t) * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(count, a.count, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int swe_set_rng(swe_ctx* c, uint64_t seed) {
  if (!c) return SWE_EINVAL;
  swe_params q = c->p;
  q.seed = seed;
  return swe_set_params(c, &q);
}

int swe_normals_noalloc(swe_ctx* c, float* out) {
  if (!c || !out) return SWE_EINVAL;
  if (c->nterr != 1) return fail(c, SWE_EINVAL, "not available on a batch context");
  if (c->ncell > kNoAllocCells)
    return fail(c, SWE_EINVAL,
                "swe_normals_noalloc: " + std::to_string(c->ncell) + " cells exceed its staging of " +
                    std::to_string(kNoAllocCells) + " cells; use swe_normals");
  float* stage = scratch_of(c->device);
  if (!stage) return fail(c, SWE_ESTATE, "swe_normals_noalloc: staging not resolved on this device");
  HIPCHK(c, hipSetDevice(c->device));
  // every plane element is written by the kernel before the copy reads it:
  // the staging needs no clearing, and a concurrent call on the same device
  // would race on it
  swe::launch_normals(c->zs[c->cur()], stage, c->nx, c->ny, c->p.mapscale, c->p.periodic != 0, c->stream);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, stage, sizeof(float) * 3 * c->ncell, hipMemcpyDeviceToHost, c->stream);
  hipError_t e2 = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(c, SWE_EHIP, std::string("normals: ") + hipGetErrorString(e));
  if (e2 != hipSuccess) return fail(c, SWE_EHIP, std::string("normals sync: ") + hipGetErrorString(e2));
  return 0;
}

}  // extern "C"
