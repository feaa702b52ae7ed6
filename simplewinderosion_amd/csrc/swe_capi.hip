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
#include <cstring>
#include <string>
#include <vector>

#include "../../include/swe.h"
#include "swe_kernels.h"

namespace {

struct Graph {
  hipGraphExec_t exec = nullptr;
  int64_t n = -1;
  int parity = -1;
  bool valid = false;
};

}  // namespace

struct swe_ctx {
  int device = 0;
  int32_t nx = 0, ny = 0;
  int64_t ncell = 0;
  hipStream_t stream = nullptr;
  swe_params p{};
  std::string err;

  // map (device)
  int64_t* zm = nullptr;
  float* zs[2] = {nullptr, nullptr};
  int64_t* dz = nullptr;
  uint32_t* vis[2] = {nullptr, nullptr};
  int64_t* sm = nullptr;
  float* ss[2] = {nullptr, nullptr};
  int64_t* ds = nullptr;
  int64_t* trk = nullptr;
  float* ema = nullptr;
  int cur = 0;  // which snapshot / event buffer the next window reads

  // particles (device)
  int64_t cap = 0;
  float4* P0 = nullptr;
  float4* P1 = nullptr;
  int64_t* L = nullptr;
  int32_t* LC = nullptr;

  // counters (device)
  int64_t* lost = nullptr;     // swe::kLostSlots
  uint64_t* dcycle = nullptr;  // RNG cycle counter
  uint64_t cycle = 0;          // host mirror

  // timing
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  struct Pending {
    hipEvent_t a, b;
    int kind;
  };
  std::vector<Pending> pending;
  swe_stats_t st{};
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
  if (p->reserved != 0) return fail(c, SWE_EINVAL, "reserved must be 0");
  if (p->max_age < -1 || p->max_age > (1 << 28)) return fail(c, SWE_EINVAL, "max_age out of range");
  if (p->periodic != 0 && p->periodic != 1) return fail(c, SWE_EINVAL, "periodic must be 0 or 1");
  if (p->track != 0 && p->track != 1) return fail(c, SWE_EINVAL, "track must be 0 or 1");
  if (p->deposit_on_death != 0 && p->deposit_on_death != 1)
    return fail(c, SWE_EINVAL, "deposit_on_death must be 0 or 1");
  return 0;
}

int ensure_layers(swe_ctx* c) {
  int r;
  if (c->p.layers == 2 && !c->sm) {
    if ((r = dalloc(c, &c->sm, c->ncell))) return r;
    if ((r = dalloc(c, &c->ss[0], c->ncell))) return r;
    if ((r = dalloc(c, &c->ss[1], c->ncell))) return r;
    if ((r = dalloc(c, &c->ds, c->ncell))) return r;
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
    if (pd.kind == 0) {
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
  p->reserved = 0;
  p->seed = 0;
}

const char* swe_version(void) { return "swe-mi355x 0.1 (gfx950)"; }

const char* swe_last_error(const swe_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int swe_destroy(swe_ctx* c) {
  if (!c) return SWE_EINVAL;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  dfree(c->zm);
  dfree(c->zs[0]);
  dfree(c->zs[1]);
  dfree(c->dz);
  dfree(c->vis[0]);
  dfree(c->vis[1]);
  dfree(c->sm);
  dfree(c->ss[0]);
  dfree(c->ss[1]);
  dfree(c->ds);
  dfree(c->trk);
  dfree(c->ema);
  dfree(c->P0);
  dfree(c->P1);
  dfree(c->L);
  dfree(c->LC);
  dfree(c->lost);
  dfree(c->dcycle);
  for (auto e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

int swe_create(swe_ctx** out, int32_t nx, int32_t ny, int32_t device) {
  if (!out) return SWE_EINVAL;
  *out = nullptr;
  if (nx < 1 || ny < 1) return SWE_EINVAL;
  if ((int64_t)nx * ny > ((int64_t)1 << 31) - 1) return SWE_EINVAL;  // cell index fits int32
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SWE_EHIP;
  if (device < 0 || device >= ndev) return SWE_EINVAL;
  swe_ctx* c = new swe_ctx();
  c->device = device;
  c->nx = nx;
  c->ny = ny;
  c->ncell = (int64_t)nx * ny;
  swe_params_default(&c->p);
  int r = 0;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    swe_destroy(c);
    return SWE_EHIP;
  }
  if ((r = dalloc(c, &c->zm, c->ncell)) || (r = dalloc(c, &c->zs[0], c->ncell)) ||
      (r = dalloc(c, &c->zs[1], c->ncell)) || (r = dalloc(c, &c->dz, c->ncell)) ||
      (r = dalloc(c, &c->vis[0], c->ncell)) || (r = dalloc(c, &c->vis[1], c->ncell)) ||
      (r = dalloc(c, &c->ema, 4 * c->ncell)) || (r = dalloc(c, &c->lost, swe::kLostSlots)) ||
      (r = dalloc(c, &c->dcycle, 1)) || (r = ensure_layers(c))) {
    swe_destroy(c);
    return r;
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess) {
    swe_destroy(c);
    return SWE_EHIP;
  }
  *out = c;
  return 0;
}

int swe_set_params(swe_ctx* c, const swe_params* p) {
  if (!c || !p) return SWE_EINVAL;
  int r = check_params(c, p);
  if (r) return r;
  HIPCHK(c, hipSetDevice(c->device));
  const bool add_layer = p->layers == 2 && c->p.layers != 2 && !c->sm;
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

int swe_upload_layers(swe_ctx* c, const float* height, const float* sediment) {
  if (!c || !height) return SWE_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  float* tmp = nullptr;
  HIPCHK(c, hipMalloc((void**)&tmp, sizeof(float) * c->ncell));
  hipError_t e = hipMemcpyAsync(tmp, height, sizeof(float) * c->ncell, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    swe::launch_from_f32(tmp, c->zm, c->zs[c->cur], c->ncell, c->stream);
    e = hipGetLastError();
  }
  if (e == hipSuccess && sediment) {
    int r = 0;
    if (!c->sm) {
      // a sediment layer can be uploaded before layers == 2 is selected
      if ((r = dalloc(c, &c->sm, c->ncell)) || (r = dalloc(c, &c->ss[0], c->ncell)) ||
          (r = dalloc(c, &c->ss[1], c->ncell)) || (r = dalloc(c, &c->ds, c->ncell))) {
        (void)hipFree(tmp);
        return r;
      }
    }
    e = hipMemcpyAsync(tmp, sediment, sizeof(float) * c->ncell, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
      swe::launch_from_f32(tmp, c->sm, c->ss[c->cur], c->ncell, c->stream);
      e = hipGetLastError();
    }
  }
  hipError_t e2 = hipStreamSynchronize(c->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return fail(c, SWE_EHIP, std::string("upload: ") + hipGetErrorString(e));
  if (e2 != hipSuccess) return fail(c, SWE_EHIP, std::string("upload sync: ") + hipGetErrorString(e2));
  return 0;
}

int swe_upload_master(swe_ctx* c, const int64_t* height_fx, const int64_t* sediment_fx) {
  if (!c) return SWE_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  if (height_fx) {
    HIPCHK(c, hipMemcpyAsync(c->zm, height_fx, sizeof(int64_t) * c->ncell, hipMemcpyHostToDevice, c->stream));
    swe::launch_snap(c->zm, c->zs[c->cur], c->ncell, c->stream);
    HIPCHK(c, hipGetLastError());
  }
  if (sediment_fx) {
    if (!c->sm) return fail(c, SWE_ESTATE, "no sediment layer: set layers = 2 first");
    HIPCHK(c, hipMemcpyAsync(c->sm, sediment_fx, sizeof(int64_t) * c->ncell, hipMemcpyHostToDevice, c->stream));
    swe::launch_snap(c->sm, c->ss[c->cur], c->ncell, c->stream);
    HIPCHK(c, hipGetLastError());
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int swe_upload_fields(swe_ctx* c, const float* massflow, const float* momentum) {
  if (!c) return SWE_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  if (massflow) HIPCHK(c, hipMemcpyAsync(c->ema, massflow, sizeof(float) * c->ncell, hipMemcpyHostToDevice, c->stream));
  if (momentum)
    HIPCHK(c, hipMemcpyAsync(c->ema + c->ncell, momentum, sizeof(float) * 3 * c->ncell, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int swe_download_layers(swe_ctx* c, float* height, float* sediment, float* massflow, float* momentum) {
  if (!c) return SWE_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  if (height) HIPCHK(c, hipMemcpyAsync(height, c->zs[c->cur], sizeof(float) * c->ncell, hipMemcpyDeviceToHost, c->stream));
  if (sediment) {
    if (c->ss[c->cur])
      HIPCHK(c, hipMemcpyAsync(sediment, c->ss[c->cur], sizeof(float) * c->ncell, hipMemcpyDeviceToHost, c->stream));
    else
      std::memset(sediment, 0, sizeof(float) * c->ncell);
  }
  if (massflow) HIPCHK(c, hipMemcpyAsync(massflow, c->ema, sizeof(float) * c->ncell, hipMemcpyDeviceToHost, c->stream));
  if (momentum)
    HIPCHK(c, hipMemcpyAsync(momentum, c->ema + c->ncell, sizeof(float) * 3 * c->ncell, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int swe_download_master(swe_ctx* c, int64_t* height_fx, int64_t* sediment_fx) {
  if (!c) return SWE_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  if (height_fx) HIPCHK(c, hipMemcpyAsync(height_fx, c->zm, sizeof(int64_t) * c->ncell, hipMemcpyDeviceToHost, c->stream));
  if (sediment_fx) {
    if (c->sm)
      HIPCHK(c, hipMemcpyAsync(sediment_fx, c->sm, sizeof(int64_t) * c->ncell, hipMemcpyDeviceToHost, c->stream));
    else
      std::memset(sediment_fx, 0, sizeof(int64_t) * c->ncell);
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int swe_erode(swe_ctx* c, int64_t particles, int32_t n_cycles) {
  if (!c) return SWE_EINVAL;
  if (particles < 0 || n_cycles < 0) return fail(c, SWE_EINVAL, "particles and n_cycles must be >= 0");
  if (particles > ((int64_t)1 << 31) - 256) return fail(c, SWE_EINVAL, "particles per cycle above 2^31");
  HIPCHK(c, hipSetDevice(c->device));
  int r;
  if ((r = ensure_layers(c))) return r;
  if ((r = ensure_particles(c, particles))) return r;

  swe::MarchArgs ma{};
  ma.n = particles;
  ma.ncell = c->ncell;
  ma.nx = c->nx;
  ma.ny = c->ny;
  ma.rcp_nx = 1.0f / (float)c->nx;
  ma.rcp_ny = 1.0f / (float)c->ny;
  ma.mapscale = c->p.mapscale;
  for (int k = 0; k < 3; k++) ma.ps[k] = c->p.pspeed[k];
  normalize3(c->p.pspeed, ma.nps);
  ma.p = c->p;
  ma.P0 = c->P0;
  ma.P1 = c->P1;
  ma.L = c->L;
  ma.LC = c->LC;
  ma.lost = c->lost;
  ma.cycle = c->dcycle;
  ma.T = c->trk;
  ma.D = c->dz;
  ma.DS = c->ds;

  swe::RelaxArgs ra{};
  ra.nx = c->nx;
  ra.ny = c->ny;
  ra.rcp_nx = ma.rcp_nx;
  ra.rcp_ny = ma.rcp_ny;
  ra.maxdiff = c->p.maxdiff;
  ra.settling = c->p.settling;
  ra.zm = c->zm;
  ra.dz = c->dz;
  ra.sm = c->sm;
  ra.ds = c->ds;

  int32_t R;
  const int64_t nwin = windows_of(c->p, &R);
  ma.R = R;
  const bool layers2 = c->p.layers == 2;

  for (int32_t cyc = 0; cyc < n_cycles; cyc++) {
    if (c->p.track) swe::launch_zero64(c->trk, 4 * c->ncell, c->stream);
    for (int64_t w = 0; w < nwin; w++) {
      const int cur = c->cur;
      ma.S = c->zs[cur];
      ma.SS = layers2 ? c->ss[cur] : nullptr;
      ma.V = c->vis[cur];
      hipEvent_t e0 = nullptr, e1 = nullptr;
      if (c->timing) {
        e0 = take_event(c);
        e1 = take_event(c);
        if (!e0 || !e1) return fail(c, SWE_EHIP, "hipEventCreate failed");
        HIPCHK(c, hipEventRecord(e0, c->stream));
      }
      swe::launch_march(ma, w == 0, c->stream);
      HIPCHK(c, hipGetLastError());
      if (c->timing) {
        HIPCHK(c, hipEventRecord(e1, c->stream));
        c->pending.push_back({e0, e1, 0});
      }
      ra.S = c->zs[cur];
      ra.SS = layers2 ? c->ss[cur] : nullptr;
      ra.V = c->vis[cur];
      ra.v_clear = c->vis[cur ^ 1];
      ra.zs_out = c->zs[cur ^ 1];
      ra.ss_out = layers2 ? c->ss[cur ^ 1] : nullptr;
      if (c->timing) {
        e0 = take_event(c);
        e1 = take_event(c);
        if (!e0 || !e1) return fail(c, SWE_EHIP, "hipEventCreate failed");
        HIPCHK(c, hipEventRecord(e0, c->stream));
      }
      swe::launch_relax(ra, c->p.periodic != 0, c->p.layers, c->stream);
      HIPCHK(c, hipGetLastError());
      if (c->timing) {
        HIPCHK(c, hipEventRecord(e1, c->stream));
        c->pending.push_back({e0, e1, 1});
      }
      c->cur ^= 1;
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

int swe_synchronize(swe_ctx* c) {
  if (!c) return SWE_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return collect_timing(c);
}

int swe_audit(swe_ctx* c, swe_audit_t* out) {
  if (!c || !out) return SWE_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  std::vector<int64_t> z((size_t)c->ncell), s;
  int64_t lost[swe::kLostSlots];
  HIPCHK(c, hipMemcpyAsync(z.data(), c->zm, sizeof(int64_t) * c->ncell, hipMemcpyDeviceToHost, c->stream));
  if (c->sm) {
    s.resize((size_t)c->ncell);
    HIPCHK(c, hipMemcpyAsync(s.data(), c->sm, sizeof(int64_t) * c->ncell, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(c, hipMemcpyAsync(lost, c->lost, sizeof(lost), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  __int128 hz = 0, hs = 0;
  for (int64_t v : z) hz += v;
  for (int64_t v : s) hs += v;
  int64_t l = 0;
  for (int k = 0; k < swe::kLostSlots; k++) l += lost[k];
  out->height_sum_lo = (int64_t)(uint64_t)hz;
  out->height_sum_hi = (int64_t)(hz >> 64);
  out->sediment_sum_lo = (int64_t)(uint64_t)hs;
  out->sediment_sum_hi = (int64_t)(hs >> 64);
  out->lost = l;
  out->cycle = c->cycle;
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

int swe_set_lost(swe_ctx* c, int64_t lost) {
  if (!c) return SWE_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  int64_t slots[swe::kLostSlots] = {0};
  slots[0] = lost;
  HIPCHK(c, hipMemcpyAsync(c->lost, slots, sizeof(slots), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int swe_set_timing(swe_ctx* c, int32_t on) {
  if (!c) return SWE_EINVAL;
  int r = collect_timing(c);
  if (r) return r;
  c->timing = on != 0;
  return 0;
}

int swe_stats(swe_ctx* c, swe_stats_t* out) {
  if (!c || !out) return SWE_EINVAL;
  int r = collect_timing(c);
  if (r) return r;
  *out = c->st;
  return 0;
}

int swe_reset_stats(swe_ctx* c) {
  if (!c) return SWE_EINVAL;
  int r = collect_timing(c);
  if (r) return r;
  std::memset(&c->st, 0, sizeof(c->st));
  return 0;
}

int swe_normals(swe_ctx* c, float* out) {
  if (!c || !out) return SWE_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  float* tmp = nullptr;
  HIPCHK(c, hipMalloc((void**)&tmp, sizeof(float) * 3 * c->ncell));
  swe::launch_normals(c->zs[c->cur], tmp, c->nx, c->ny, c->p.mapscale, c->p.periodic != 0, c->stream);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, tmp, sizeof(float) * 3 * c->ncell, hipMemcpyDeviceToHost, c->stream);
  hipError_t e2 = hipStreamSynchronize(c->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return fail(c, SWE_EHIP, std::string("normals: ") + hipGetErrorString(e));
  if (e2 != hipSuccess) return fail(c, SWE_EHIP, std::string("normals sync: ") + hipGetErrorString(e2));
  return 0;
}

}  // extern "C"
