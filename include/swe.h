/*
 * swe.h -- C-ABI of the MI355X (gfx950) erosion library, libswe.so.
 *
 * The library replaces SimpleWindErosion's erosion hot path:
 *   World::erode(int cycles)   source/world.hpp:21,36-77
 *   Wind::Wind(vec2), fly()    source/wind.hpp:6,20,28-117
 *   World::cascade(vec2)       source/world.hpp:22,79-153
 *   quad::map get/height/normal, quad::_normal
 *                              source/cellpool.h:200-223,432-480
 * Plain C types only; every call returns 0 on success or a negative
 * swe_status, and swe_last_error() describes the last failure of a context.
 *
 * Semantics (batched-synchronous): within one snapshot window every particle
 * makes `window` fly() calls reading the same float32 snapshot of the map;
 * pickups are summed as int64 fixed point with 2^-32 resolution.  Each call
 * counts one cascade event at its pre-move cell; at the end of the window
 * every cell with V > 0 events runs World::cascade once against the
 * snapshot (8 neighbours in the reference's offset order, stable sort by
 * height, walk with the centre evolving) with the effective settling factor
 * 1 - (1 - settling)^V -- what V consecutive events leave of an isolated
 * pair's excess -- and the net transfers are summed in the same fixed
 * point.  The default window of 1 makes the method step-synchronous: every
 * particle takes one step against the same map, then the map is updated.
 * Results are deterministic for given (params, seed, cycle counter) and
 * independent of thread and launch order.
 *
 * Stability: where the neighbourhoods of cells visited in one window
 * overlap, shared edges move once per cascade that touches them, so large
 * settling factors overshoot.  settling <= 0.1 is stable on every workload
 * measured (the benchmark workloads use 0.01); swe_set_params rejects
 * settling > 0.25 with the synchronous cascade (cascade_order 0); the
 * colour-ordered cascade (cascade_order 1, or -1 above 0.25) runs the
 * reference's own settling = 1.0.  Independently of the cascade, pickups of
 * many particles in one cell in one window add up: with the reference's
 * suspension 0.05 the method stays bounded at the reference's density
 * (4096 particles on 512^2) but not at one particle per cell, where the
 * benchmark workloads use suspension 1e-4.
 *
 * Map layout: cell (x, y) lives at index x*ny + y (math::flatten,
 * source/include/math.h:11-14): x is the slow axis, y the fast axis.
 * The reference's third axis (pos.z) is the map's y axis.
 *
 * Threading: one context per thread; calls on a context are serialized on
 * its HIP stream.  Two contexts may share a device.
 */
#ifndef SWE_H
#define SWE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct swe_ctx swe_ctx;

typedef enum {
  SWE_OK = 0,
  SWE_EINVAL = -1,      /* bad argument */
  SWE_EHIP = -2,        /* HIP runtime error */
  SWE_ENOMEM = -3,      /* device allocation failed */
  SWE_EAUDIT = -4,      /* mass audit mismatch */
  SWE_ESTATE = -5       /* call not valid in the current state */
} swe_status;

/*
 * Parameters.  Replaces the mutable statics Wind::maxAge/boundarylayer/
 * suspension/gravity (wind.hpp:15-26), Wind::pspeed (wind.hpp:13) and
 * World::lrate/maxdiff/settling (world.hpp:17-19,30-32), plus the
 * compile-time quad::mapscale (cellpool.h:184) and the boundary choice
 * (bounded cellpool.h:82-94,432-445 or periodic cellpool.h:96-113,447-465).
 * quad::lodsize is fixed at 1 as in the reference.
 */
typedef struct swe_params {
  int32_t max_age;        /* Wind::maxAge = 512 -> 513 particle-steps */
  float boundarylayer;    /* Wind::boundarylayer = 2.0 */
  float suspension;       /* Wind::suspension = 0.05 */
  float gravity;          /* Wind::gravity = 0.1 */
  float pspeed[3];        /* Wind::pspeed = normalize(vec3(1,0,0)) */
  float lrate;            /* World::lrate = 0.1 */
  float maxdiff;          /* World::maxdiff = 0.005 */
  float settling;         /* World::settling = 1.0 */
  float abrasion;         /* sediment-layer extension (layers == 2): bedrock pickup rate */
  float mapscale;         /* quad::mapscale = 80 */
  int32_t window;         /* fly() calls per snapshot window (default 1); 0 = one window per cycle */
  int32_t periodic;       /* 1 = periodic map, 0 = bounded (reference default) */
  int32_t track;          /* 1 = massflow/momentum tracking and EMA fields */
  int32_t layers;         /* 1 = height only (reference); 2 = height + loose sediment */
  int32_t deposit_on_death; /* 1 = a dying particle drops its load (closes map mass) */
  int32_t spawn;          /* 0 = uniform over the map (world.hpp:52-54), 1 = on the x = 0 / y = 0
                             edges (the commented variant, world.hpp:56-61) */
  uint64_t seed;          /* World::SEED: feeds the counter RNG */
  int32_t cascade_order;  /* -1 = automatic (default): colour-ordered when settling > 0.25, else
                             synchronous; 0 = synchronous (every visited cell's cascade reads the
                             window's snapshot; settling <= 0.25 only); 1 = colour-ordered (cells in
                             3x3 colour classes, one class after another, each reading the map the
                             earlier classes left: the reference's sequential cascades in class
                             order; global path, single maps) */
  int32_t reserved;       /* 0 */
} swe_params;

/* Mass audit: exact in int64 fixed point (2^-32). */
typedef struct swe_audit_t {
  int64_t height_sum_lo;  /* sum over cells of height, low 64 bits of an int128 */
  int64_t height_sum_hi;  /* high 64 bits */
  int64_t sediment_sum_lo;/* layers == 2 */
  int64_t sediment_sum_hi;
  int64_t lost;           /* mass carried off by particles that died loaded */
  uint64_t cycle;         /* cycles completed (RNG counter) */
} swe_audit_t;

typedef struct swe_stats_t {
  uint64_t cycles;        /* map-cycles run on this context */
  uint64_t particle_steps;/* fly() calls that returned true */
  double march_ms;        /* HIP-event time of the particle march kernels (timing on) */
  uint64_t march_launches;
  double relax_ms;        /* HIP-event time of the relax kernels (timing on) */
  uint64_t relax_launches;
  uint64_t max_tile_particles;  /* most particles one step-kernel tile held in a window */
  uint64_t crowded_tile_windows;/* (tile, window) pairs above the tile's LDS list capacity */
  uint64_t far_particles;       /* particle moves that skipped more than one tile */
  uint64_t heavy_tile_windows;  /* (tile, window) pairs run first as crowded (SWE_TUNE_HEAVY_FIRST) */
  uint64_t light_tile_windows;  /* (tile, window) pairs run last as light (SWE_TUNE_HEAVY_FIRST = 2) */
} swe_stats_t;

/* Reference defaults (wind.hpp:13,23-26; world.hpp:30-32; cellpool.h:184). */
void swe_params_default(swe_params* p);

/*
 * Create a context owning an nx x ny map on HIP device `device`.
 * Replaces World::map / mappool::pool<cell> (world.hpp:13,28;
 * SimpleWindErosion.cpp:11,36; cellpool.h:139-162,335-356).
 * The map starts flat (height 0) with reference default parameters.
 */
int swe_create(swe_ctx** out, int32_t nx, int32_t ny, int32_t device);
int swe_destroy(swe_ctx* ctx);
const char* swe_last_error(const swe_ctx* ctx);

int swe_set_params(swe_ctx* ctx, const swe_params* p);
int swe_get_params(const swe_ctx* ctx, swe_params* p);

/*
 * Host <-> device map transfer (caller-owned host memory, nx*ny floats,
 * index x*ny + y).  Upload converts to the fixed-point master layers and
 * refreshes the snapshot; sediment may be NULL (zero thickness).
 * Download returns the current snapshot of each requested layer; any
 * pointer may be NULL.  momentum is 3*nx*ny floats (x, y, z planes).
 */
int swe_upload_layers(swe_ctx* ctx, const float* height, const float* sediment);
int swe_download_layers(swe_ctx* ctx, float* height, float* sediment, float* massflow,
                        float* momentum);
/* Raw int64 fixed-point master layers (exact state), nx*ny each; NULL skips. */
int swe_download_master(swe_ctx* ctx, int64_t* height_fx, int64_t* sediment_fx);
int swe_upload_master(swe_ctx* ctx, const int64_t* height_fx, const int64_t* sediment_fx);
/* EMA flow fields in: massflow nx*ny, momentum 3*nx*ny; NULL skips. */
int swe_upload_fields(swe_ctx* ctx, const float* massflow, const float* momentum);

/*
 * World::erode(int cycles) (world.hpp:36-77): run n_cycles map-cycles, each
 * spawning `particles_per_cycle` particles (the reference's `cycles`
 * argument) and marching every one of them to death.  Asynchronous on the
 * context's stream; swe_synchronize() waits.
 */
int swe_erode(swe_ctx* ctx, int64_t particles_per_cycle, int32_t n_cycles);
int swe_synchronize(swe_ctx* ctx);

/*
 * Execution path of swe_erode (results are identical on every path):
 * 0 = automatic (tile-owned step kernel for window == 1, else the global
 * march), 1 = global-memory march with device-scope atomics (any window),
 * 2 = tile-owned step kernel (window == 1 only).  No reference counterpart:
 * the reference has a single sequential path.
 */
int swe_set_path(swe_ctx* ctx, int32_t path);

/*
 * Launch tuning knobs: they change speed only, never results.  No reference
 * counterpart.
 *   SWE_TUNE_TILE_ORDER: order in which the step kernel's workgroups take
 *     tiles; 0 (default) = each of the 8 XCDs a contiguous range of tiles,
 *     C in [1, 4096] = runs of C consecutive tiles dealt round-robin over the
 *     XCDs.
 *   SWE_TUNE_GRAPH: 1 (default) = the tile path captures a map-cycle's
 *     kernel launches into a HIP graph once (per starting window parity and
 *     launch configuration) and launches the graph each cycle; 0 = launches
 *     each kernel.  Timing (swe_set_timing) uses individual launches.
 *   SWE_TUNE_HEAVY_FIRST: 1 = a tile that received more than 1024 particles
 *     in one window is handed to one of the first workgroups of the next
 *     window's launch, so crowded tiles start first instead of finishing
 *     last; 2 = also a tile that received at most 256 is handed to one of
 *     the last workgroups (on its own XCD), so short workgroups fill the
 *     launch's tail; 0 = every tile in the SWE_TUNE_TILE_ORDER order.
 *     Default: 2 for contexts of up to 4096 tiles (16x32 cells each), 1
 *     above (measured best on a 1024^2 map and on a 64-terrain batch).
 *     Contexts of fewer than 64 tiles always use 0.
 *   SWE_TUNE_SHAPE: the step kernel's tile shape, an index into the
 *     compiled shapes 0 = 16x32 cells / 256 threads / 5 waves per SIMD
 *     (default), 1 = 32x16 / 256 / 5, 2 = 16x16 / 128 / 5, 3 = 16x32 / 256 / 4,
 *     4 = 32x32 / 512 / 4 (x the slow axis).  Changing it synchronizes the
 *     context's stream and rebuilds its per-tile buffers.
 */
enum swe_tune_knob { SWE_TUNE_TILE_ORDER = 1, SWE_TUNE_GRAPH = 2, SWE_TUNE_HEAVY_FIRST = 3, SWE_TUNE_SHAPE = 4 };
int swe_set_tune(swe_ctx* ctx, int32_t knob, int64_t value);
int swe_get_tune(swe_ctx* ctx, int32_t knob, int64_t* value);

/*
 * Autotuner (no reference counterpart): times `trial_cycles` map-cycles of
 * the context's own workload (particles per terrain, as for
 * swe_erode_batch), each from the same saved state: first every tile shape
 * (SWE_TUNE_SHAPE) with graphs, then on the fastest shape every tile order
 * in {0, 1, 4, 16} with and without graphs (and, on maps of 64 tiles or
 * more, every SWE_TUNE_HEAVY_FIRST mode); then restores the state (map,
 * fields, counters and statistics are as before the call) and keeps the
 * fastest setting.  Tile step path (window == 1) only.
 */
typedef struct swe_tune_result {
  int32_t tile_order;
  int32_t graph;
  int32_t candidates;    /* settings timed */
  int32_t heavy_first;
  double ms_per_cycle;   /* of the setting kept */
  int32_t shape;         /* SWE_TUNE_SHAPE kept */
  int32_t reserved;
} swe_tune_result;
int swe_autotune(swe_ctx* ctx, const int64_t* particles, int32_t trial_cycles, swe_tune_result* out);
/* The GPU architecture of the context's device (e.g. "gfx950:sramecc+:xnack-"), the key of tuning caches. */
int swe_device_arch(swe_ctx* ctx, char* buf, int32_t len);

/*
 * Batches: many terrains of different sizes in one context, eroded together
 * by one launch per snapshot window (window == 1).  Terrain k's cells, tiles
 * and particles are contiguous ranges of the context's buffers; results of
 * each terrain equal those of a single-terrain context with the same size,
 * inputs, seed and wind.  Shared parameters come from swe_set_params; each
 * terrain may override the seed and the prevailing wind (pspeed may be
 * NULL).  Single-map calls (swe_upload_layers, swe_erode, ...) act on a
 * context of one terrain; batch contexts use the *_terrain / *_batch calls.
 * No reference counterpart: the reference erodes one map.
 */
int swe_create_batch(swe_ctx** out, int32_t n_terrains, const int32_t* nx, const int32_t* ny, int32_t device);
int swe_get_terrains(const swe_ctx* ctx, int32_t* n_terrains);
int swe_set_terrain(swe_ctx* ctx, int32_t k, uint64_t seed, const float* pspeed);
int swe_upload_terrain(swe_ctx* ctx, int32_t k, const float* height, const float* sediment);
int swe_download_terrain(swe_ctx* ctx, int32_t k, float* height, float* sediment, float* massflow, float* momentum);
int swe_download_terrain_master(swe_ctx* ctx, int32_t k, int64_t* height_fx, int64_t* sediment_fx);
int swe_audit_terrain(swe_ctx* ctx, int32_t k, struct swe_audit_t* out);
/* One map-cycle per n_cycles for every terrain; particles[k] per terrain. */
int swe_erode_batch(swe_ctx* ctx, const int64_t* particles, int32_t n_cycles);
/*
 * Terrain k's size, effective seed and prevailing wind (the context's when
 * not overridden); overrides: bit 0 = seed set by swe_set_terrain, bit 1 =
 * wind set.  Any output pointer may be NULL.
 */
int swe_get_terrain(const swe_ctx* ctx, int32_t k, int32_t* nx, int32_t* ny, uint64_t* seed, float* pspeed,
                    int32_t* overrides);
/* Batch counterparts of swe_upload_master, swe_upload_fields and swe_set_lost. */
int swe_upload_terrain_master(swe_ctx* ctx, int32_t k, const int64_t* height_fx, const int64_t* sediment_fx);
int swe_upload_terrain_fields(swe_ctx* ctx, int32_t k, const float* massflow, const float* momentum);
int swe_set_terrain_lost(swe_ctx* ctx, int32_t k, int64_t lost);
/* Batch counterparts of swe_windmap and swe_normals (same layouts, terrain k's cells). */
int swe_windmap_terrain(swe_ctx* ctx, int32_t k, float* out);
int swe_normals_terrain(swe_ctx* ctx, int32_t k, float* out);

/* Exact mass audit of the current state (synchronizes). */
int swe_audit(swe_ctx* ctx, swe_audit_t* out);
/* RNG cycle counter (for resume); setting it does not touch the map. */
int swe_get_cycle(const swe_ctx* ctx, uint64_t* cycle);
int swe_set_cycle(swe_ctx* ctx, uint64_t cycle);
/* Carried-off mass counter (fixed point) for resume. */
int swe_set_lost(swe_ctx* ctx, int64_t lost);

/* Map size of a context. */
int swe_get_size(const swe_ctx* ctx, int32_t* nx, int32_t* ny);

/*
 * Checkpoint / resume: the exact int64 master layers, EMA fields, params
 * (as JSON text), cycle counter, and per terrain its seed, wind and
 * lost-mass counter, in one little-endian file, so that load + erode
 * continues bit for bit; single maps and batches (a checkpoint loads into a
 * context of the same terrain sizes).  No counterpart in the reference,
 * whose only state is its seed (SimpleWindErosion.cpp:27-34).
 */
int swe_save_state(swe_ctx* ctx, const char* path);
int swe_load_state(swe_ctx* ctx, const char* path);

/*
 * Kernel timing with HIP events on the context stream (off by default).
 * on = 0: off; on = n >= 1: events around every n-th launch of the march /
 * relax / step kernels (n > 1 samples the launches, leaving the others
 * without the events' extra queue packets).
 */
int swe_set_timing(swe_ctx* ctx, int32_t on);
int swe_stats(swe_ctx* ctx, swe_stats_t* out);
int swe_reset_stats(swe_ctx* ctx);

/*
 * The wind map of the reference's map view (SimpleWindErosion.cpp:224-230):
 * per cell (0.5(1+erf(2 mx)), 0.5(1+erf(2 my)), 0.5(1+erf(2 mz)), 1) from the
 * EMA momentum field.  out: 4*nx*ny floats, RGBA interleaved per cell.
 */
int swe_windmap(swe_ctx* ctx, float* out);

/*
 * Flow-field and map products written to a file (the reference draws its
 * wind map into a texture each frame, SimpleWindErosion.cpp:224-230, and
 * writes no files).  SWE_EXPORT_NPZ: a numpy archive with height, sediment,
 * massflow (nx, ny), momentum and normals (3, nx, ny), windmap (nx, ny, 4),
 * all float32, and the exact masters height_fx, sediment_fx (nx, ny) int64.
 * SWE_EXPORT_PNG: the wind map as 8-bit RGBA, nx wide and ny high, pixel
 * (x, y) = cell (x, y), channels round(255 v).
 */
enum swe_export_format { SWE_EXPORT_NPZ = 0, SWE_EXPORT_PNG = 1 };
int swe_export(swe_ctx* ctx, const char* path, int32_t format);
int swe_export_terrain(swe_ctx* ctx, int32_t k, const char* path, int32_t format);

/*
 * Render-mesh vertex attributes as updatenode writes them (cellpool.h:302-333):
 * per cell P = (x, mapscale*h, y), normal, T - P, B - P (T, B the +x / +y
 * neighbours; height 0 off a bounded map).  out: 12*nx*ny floats per cell
 * in that order.
 */
int swe_mesh(swe_ctx* ctx, float* out);

/*
 * Surface normals of the current snapshot, quad::_normal (cellpool.h:200-223)
 * for every cell, as used by updatenode (cellpool.h:302-333).  out: 3*nx*ny
 * floats, planes x, y, z.
 */
int swe_normals(swe_ctx* ctx, float* out);

/*
 * swe_normals without any device allocation: the same kernel writes into a
 * staging array the library holds in device memory for the life of the
 * process (one per device, 12 MiB: maps of at most 1,048,576 cells; larger
 * maps return SWE_EINVAL).  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (per-frame render loops,
 * serving loops, code that forbids hipMalloc).  Calls that overlap on one
 * device race on the staging array.
 */
int swe_normals_noalloc(swe_ctx* ctx, float* out);

/*
 * Known-answer entry points (not on the hot path; a map context only).
 *
 * swe_debug_step: one Wind::fly() call (wind.hpp:28-117, with the cell
 * lookup of :30-38) of each of n caller-given particles against the current
 * snapshot, by the device code the step kernels run.  Host arrays:
 *   state  n x 7 floats in/out: pos x, y, z, speed x, y, z, sediment
 *   age    n in/out: calls made (0 = fresh: pos.y snaps to the ground);
 *          -1 once the particle died
 *   key    n RNG keys: the turbulence draws of call `age` (wind.hpp:83)
 *   alive  n out: 1 for a particle-step, 0 when the particle died (left a
 *          bounded map or passed max_age)
 *   pickup n out: the fixed-point (2^-32) amount the call took from the
 *          particle's cell (wind.hpp:110-112; negative: deposited)
 *
 * swe_debug_cascade: one World::cascade (world.hpp:79-153) at cell (x, y)
 * against the current snapshot, with the settling factor of `events`
 * cascade events (1 - (1 - settling)^events): count = in-map neighbours,
 * nb_cell[8] their cells in the reference's offset order (world.hpp:83-92)
 * and nb_fx[8] the fixed-point amount each receives (the centre loses the
 * sum).
 */
int swe_debug_step(swe_ctx* ctx, int32_t n, float* state, int32_t* age, const uint64_t* key, int32_t* alive,
                   int64_t* pickup);
int swe_debug_cascade(swe_ctx* ctx, int32_t x, int32_t y, uint32_t events, int32_t* count, int64_t* nb_cell,
                      int64_t* nb_fx);

/*
 * The particles' RNG seed (World::SEED, SimpleWindErosion.cpp:27-34, which
 * seeds srand): the same as swe_set_params with params.seed = seed.
 */
int swe_set_rng(swe_ctx* ctx, uint64_t seed);

/*
 * Terrain generator for the benchmark inputs: seeded fBm gradient noise
 * (hashed lattice gradients, quintic fade, `octaves` octaves with lacunarity
 * 2 and gain 0.5, base frequency `base_cycles` periods across the map,
 * periodic in both axes) scaled to [0,1], plus `n_pyramids` square pyramids
 * with heights in [pyr_hmin, pyr_hmax] and base widths in [pyr_bmin,
 * pyr_bmax] cells at seeded positions.  Host-side; out: nx*ny floats.
 */
int swe_generate_terrain(int32_t nx, int32_t ny, uint64_t seed, int32_t octaves,
                         float base_cycles, int32_t n_pyramids, float pyr_hmin,
                         float pyr_hmax, float pyr_bmin, float pyr_bmax, float* out);

/* Library version string. */
const char* swe_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SWE_H */
