// swe_state.cpp -- checkpoint / resume of a context's exact state.
//
// The reference keeps no state beyond its seed (SimpleWindErosion.cpp:27-34).
// A checkpoint here is the exact int64 master layers plus everything the
// next cycle depends on, so that load + erode continues bit for bit, for a
// single map or a batch of terrains.  Format 2 (written):
//
//   offset  size     field
//   0       8        magic "SWESTAT2"
//   8       8        L: length of the header text (uint64, little endian)
//   16      L        header, JSON: {"format": "swe-state", "version": 2,
//                    "cycle": C, "params": {every swe_params field by name},
//                    "terrains": [{"nx", "ny", "seed", "pspeed", "overrides",
//                    "lost"}, ...]}; floats printed with 9 significant
//                    digits (exact for float32), integers in decimal
//   16+L    ...      per terrain, in order: height master int64[nx*ny]
//                    (fixed point 2^-32); sediment master int64[nx*ny] when
//                    layers == 2; EMA fields massflow, momentum x, y, z as
//                    float32[nx*ny] each
//
// Format 1 ("SWESTAT1", single maps, parameters as the raw 80 bytes of
// swe_params through `seed`) is still read, with the synchronous cascade.
#include <cctype>
#include <cinttypes>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/swe.h"

namespace {
const char kMagic1[8] = {'S', 'W', 'E', 'S', 'T', 'A', 'T', '1'};
const char kMagic2[8] = {'S', 'W', 'E', 'S', 'T', 'A', 'T', '2'};
static_assert(sizeof(swe_params) == 88, "swe_params layout");
// format 1 stored the first 80 bytes of swe_params (through seed)
constexpr size_t kParamsV1 = 80;

// ---- header text ---------------------------------------------------------
std::string fmt_float(float v) {
  char b[32];
  std::snprintf(b, sizeof b, "%.9g", (double)v);
  return b;
}
std::string fmt_i64(int64_t v) {
  char b[32];
  std::snprintf(b, sizeof b, "%" PRId64, v);
  return b;
}
std::string fmt_u64(uint64_t v) {
  char b[32];
  std::snprintf(b, sizeof b, "%" PRIu64, v);
  return b;
}

std::string params_json(const swe_params& p) {
  std::string s = "{";
  s += "\"max_age\": " + fmt_i64(p.max_age);
  s += ", \"boundarylayer\": " + fmt_float(p.boundarylayer);
  s += ", \"suspension\": " + fmt_float(p.suspension);
  s += ", \"gravity\": " + fmt_float(p.gravity);
  s += ", \"pspeed\": [" + fmt_float(p.pspeed[0]) + ", " + fmt_float(p.pspeed[1]) + ", " + fmt_float(p.pspeed[2]) + "]";
  s += ", \"lrate\": " + fmt_float(p.lrate);
  s += ", \"maxdiff\": " + fmt_float(p.maxdiff);
  s += ", \"settling\": " + fmt_float(p.settling);
  s += ", \"abrasion\": " + fmt_float(p.abrasion);
  s += ", \"mapscale\": " + fmt_float(p.mapscale);
  s += ", \"window\": " + fmt_i64(p.window);
  s += ", \"periodic\": " + fmt_i64(p.periodic);
  s += ", \"track\": " + fmt_i64(p.track);
  s += ", \"layers\": " + fmt_i64(p.layers);
  s += ", \"deposit_on_death\": " + fmt_i64(p.deposit_on_death);
  s += ", \"spawn\": " + fmt_i64(p.spawn);
  s += ", \"seed\": " + fmt_u64(p.seed);
  s += ", \"cascade_order\": " + fmt_i64(p.cascade_order);
  s += "}";
  return s;
}

// A JSON value of the header: numbers are kept as their text.
struct JVal {
  enum Kind { kNull, kNum, kStr, kArr, kObj } kind = kNull;
  std::string text;                             // kNum, kStr
  std::vector<JVal> arr;                        // kArr
  std::vector<std::pair<std::string, JVal>> obj;  // kObj
  const JVal* get(const char* key) const {
    for (auto& kv : obj)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
};

struct JParser {
  const char* p;
  const char* end;
  bool ok = true;
  void ws() {
    while (p < end && std::isspace((unsigned char)*p)) p++;
  }
  bool eat(char ch) {
    ws();
    if (p < end && *p == ch) {
      p++;
      return true;
    }
    return false;
  }
  JVal value(int depth) {
    JVal v;
    ws();
    if (depth > 8 || p >= end) {
      ok = false;
      return v;
    }
    if (*p == '{') {
      p++;
      v.kind = JVal::kObj;
      if (eat('}')) return v;
      do {
        JVal k = value(depth + 1);
        if (!ok || k.kind != JVal::kStr || !eat(':')) {
          ok = false;
          return v;
        }
        JVal x = value(depth + 1);
        if (!ok) return v;
        v.obj.emplace_back(k.text, std::move(x));
      } while (eat(','));
      if (!eat('}')) ok = false;
    } else if (*p == '[') {
      p++;
      v.kind = JVal::kArr;
      if (eat(']')) return v;
      do {
        v.arr.push_back(value(depth + 1));
        if (!ok) return v;
      } while (eat(','));
      if (!eat(']')) ok = false;
    } else if (*p == '"') {
      p++;
      v.kind = JVal::kStr;
      while (p < end && *p != '"') v.text += *p++;
      if (p >= end) ok = false;
      else p++;
    } else {
      v.kind = JVal::kNum;
      while (p < end && (std::isdigit((unsigned char)*p) || *p == '-' || *p == '+' || *p == '.' || *p == 'e' ||
                         *p == 'E' || *p == 'i' || *p == 'n' || *p == 'f' || *p == 'a'))
        v.text += *p++;
      if (v.text.empty()) ok = false;
    }
    return v;
  }
};

bool num_i64(const JVal* v, int64_t* out) {
  if (!v || v->kind != JVal::kNum) return false;
  char* e = nullptr;
  const long long x = std::strtoll(v->text.c_str(), &e, 10);
  if (*e) return false;
  *out = (int64_t)x;
  return true;
}
bool num_u64(const JVal* v, uint64_t* out) {
  if (!v || v->kind != JVal::kNum || v->text[0] == '-') return false;
  char* e = nullptr;
  const unsigned long long x = std::strtoull(v->text.c_str(), &e, 10);
  if (*e) return false;
  *out = (uint64_t)x;
  return true;
}
bool num_i32(const JVal* v, int32_t* out) {
  int64_t x;
  if (!num_i64(v, &x) || x < INT32_MIN || x > INT32_MAX) return false;
  *out = (int32_t)x;
  return true;
}
bool num_f32(const JVal* v, float* out) {
  if (!v || v->kind != JVal::kNum) return false;
  char* e = nullptr;
  const float x = std::strtof(v->text.c_str(), &e);
  if (*e) return false;
  *out = x;
  return true;
}
bool vec3_f32(const JVal* v, float out[3]) {
  if (!v || v->kind != JVal::kArr || v->arr.size() != 3) return false;
  for (int q = 0; q < 3; q++)
    if (!num_f32(&v->arr[(size_t)q], &out[q])) return false;
  return true;
}

bool params_from_json(const JVal* j, swe_params* p) {
  if (!j || j->kind != JVal::kObj) return false;
  swe_params_default(p);
  return num_i32(j->get("max_age"), &p->max_age) && num_f32(j->get("boundarylayer"), &p->boundarylayer) &&
         num_f32(j->get("suspension"), &p->suspension) && num_f32(j->get("gravity"), &p->gravity) &&
         vec3_f32(j->get("pspeed"), p->pspeed) && num_f32(j->get("lrate"), &p->lrate) &&
         num_f32(j->get("maxdiff"), &p->maxdiff) && num_f32(j->get("settling"), &p->settling) &&
         num_f32(j->get("abrasion"), &p->abrasion) && num_f32(j->get("mapscale"), &p->mapscale) &&
         num_i32(j->get("window"), &p->window) && num_i32(j->get("periodic"), &p->periodic) &&
         num_i32(j->get("track"), &p->track) && num_i32(j->get("layers"), &p->layers) &&
         num_i32(j->get("deposit_on_death"), &p->deposit_on_death) && num_i32(j->get("spawn"), &p->spawn) &&
         num_u64(j->get("seed"), &p->seed) && num_i32(j->get("cascade_order"), &p->cascade_order);
}

struct File {
  FILE* f;
  explicit File(FILE* f_) : f(f_) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

// format 1: one map, raw parameter struct
int load_v1(swe_ctx* ctx, FILE* f) {
  int32_t nx = 0, ny = 0, fnx = 0, fny = 0;
  int r;
  if ((r = swe_get_size(ctx, &nx, &ny))) return r;
  swe_params p;
  swe_params_default(&p);
  p.cascade_order = 0;  // format 1 predates the colour-ordered cascade
  uint64_t cycle = 0;
  int64_t lost = 0;
  bool ok = std::fread(&fnx, 4, 1, f) == 1 && std::fread(&fny, 4, 1, f) == 1 && fnx == nx && fny == ny &&
            std::fread(&p, kParamsV1, 1, f) == 1 && std::fread(&cycle, 8, 1, f) == 1 &&
            std::fread(&lost, 8, 1, f) == 1;
  const size_t n = (size_t)nx * ny;
  std::vector<int64_t> h(n), s;
  std::vector<float> mf(n), mo(3 * n);
  if (ok) ok = std::fread(h.data(), 8, n, f) == n;
  if (ok && p.layers == 2) {
    s.resize(n);
    ok = std::fread(s.data(), 8, n, f) == n;
  }
  if (ok) ok = std::fread(mf.data(), 4, n, f) == n && std::fread(mo.data(), 4, 3 * n, f) == 3 * n;
  if (!ok) return SWE_EINVAL;
  if ((r = swe_set_params(ctx, &p))) return r;
  if ((r = swe_upload_master(ctx, h.data(), p.layers == 2 ? s.data() : nullptr))) return r;
  if ((r = swe_upload_fields(ctx, mf.data(), mo.data()))) return r;
  if ((r = swe_set_cycle(ctx, cycle))) return r;
  return swe_set_lost(ctx, lost);
}
}  // namespace

extern "C" int swe_save_state(swe_ctx* ctx, const char* path) {
  if (!ctx || !path) return SWE_EINVAL;
  swe_params p;
  int r;
  int32_t nterr = 0;
  if ((r = swe_get_params(ctx, &p)) || (r = swe_get_terrains(ctx, &nterr))) return r;
  uint64_t cycle = 0;
  if ((r = swe_get_cycle(ctx, &cycle))) return r;
  std::string hdr = "{\"format\": \"swe-state\", \"version\": 2, \"cycle\": " + fmt_u64(cycle) +
                    ", \"params\": " + params_json(p) + ", \"terrains\": [";
  std::vector<std::pair<int32_t, int32_t>> sizes;
  for (int32_t k = 0; k < nterr; k++) {
    int32_t nx = 0, ny = 0, ov = 0;
    uint64_t seed = 0;
    float ps[3];
    if ((r = swe_get_terrain(ctx, k, &nx, &ny, &seed, ps, &ov))) return r;
    swe_audit_t a;
    if ((r = swe_audit_terrain(ctx, k, &a))) return r;
    sizes.emplace_back(nx, ny);
    hdr += std::string(k ? ", " : "") + "{\"nx\": " + fmt_i64(nx) + ", \"ny\": " + fmt_i64(ny) +
           ", \"seed\": " + fmt_u64(seed) + ", \"pspeed\": [" + fmt_float(ps[0]) + ", " + fmt_float(ps[1]) + ", " +
           fmt_float(ps[2]) + "], \"overrides\": " + fmt_i64(ov) + ", \"lost\": " + fmt_i64(a.lost) + "}";
  }
  hdr += "]}\n";
  File f(std::fopen(path, "wb"));
  if (!f.f) return SWE_EINVAL;
  const uint64_t L = hdr.size();
  bool ok = std::fwrite(kMagic2, 1, 8, f.f) == 8 && std::fwrite(&L, 8, 1, f.f) == 1 &&
            std::fwrite(hdr.data(), 1, hdr.size(), f.f) == hdr.size();
  for (int32_t k = 0; ok && k < nterr; k++) {
    const size_t n = (size_t)sizes[(size_t)k].first * sizes[(size_t)k].second;
    std::vector<int64_t> h(n), s(n);
    std::vector<float> mf(n), mo(3 * n);
    if ((r = swe_download_terrain_master(ctx, k, h.data(), s.data()))) return r;
    if ((r = swe_download_terrain(ctx, k, nullptr, nullptr, mf.data(), mo.data()))) return r;
    ok = std::fwrite(h.data(), 8, n, f.f) == n;
    if (ok && p.layers == 2) ok = std::fwrite(s.data(), 8, n, f.f) == n;
    if (ok) ok = std::fwrite(mf.data(), 4, n, f.f) == n && std::fwrite(mo.data(), 4, 3 * n, f.f) == 3 * n;
  }
  FILE* fp = f.f;
  f.f = nullptr;
  ok = (std::fclose(fp) == 0) && ok;
  return ok ? 0 : SWE_EINVAL;
}

extern "C" int swe_load_state(swe_ctx* ctx, const char* path) {
  if (!ctx || !path) return SWE_EINVAL;
  File f(std::fopen(path, "rb"));
  if (!f.f) return SWE_EINVAL;
  char magic[8];
  if (std::fread(magic, 1, 8, f.f) != 8) return SWE_EINVAL;
  if (std::memcmp(magic, kMagic1, 8) == 0) {
    int32_t nterr = 0;
    if (swe_get_terrains(ctx, &nterr) || nterr != 1) return SWE_EINVAL;
    return load_v1(ctx, f.f);
  }
  if (std::memcmp(magic, kMagic2, 8) != 0) return SWE_EINVAL;
  uint64_t L = 0;
  if (std::fread(&L, 8, 1, f.f) != 1 || L > (1u << 26)) return SWE_EINVAL;
  std::string hdr((size_t)L, '\0');
  if (std::fread(&hdr[0], 1, (size_t)L, f.f) != (size_t)L) return SWE_EINVAL;
  JParser jp{hdr.data(), hdr.data() + hdr.size()};
  const JVal root = jp.value(0);
  if (!jp.ok || root.kind != JVal::kObj) return SWE_EINVAL;
  int64_t version = 0;
  uint64_t cycle = 0;
  swe_params p;
  const JVal* terr = root.get("terrains");
  if (!num_i64(root.get("version"), &version) || version != 2 || !num_u64(root.get("cycle"), &cycle) ||
      !params_from_json(root.get("params"), &p) || !terr || terr->kind != JVal::kArr)
    return SWE_EINVAL;
  int32_t nterr = 0;
  int r;
  if ((r = swe_get_terrains(ctx, &nterr))) return r;
  if ((size_t)nterr != terr->arr.size()) return SWE_EINVAL;
  // every terrain's geometry must be the context's before anything changes
  struct T {
    int32_t nx, ny, ov;
    uint64_t seed;
    float ps[3];
    int64_t lost;
  };
  std::vector<T> ts((size_t)nterr);
  for (int32_t k = 0; k < nterr; k++) {
    const JVal& t = terr->arr[(size_t)k];
    T& d = ts[(size_t)k];
    int64_t ov = 0;
    if (!num_i32(t.get("nx"), &d.nx) || !num_i32(t.get("ny"), &d.ny) || !num_u64(t.get("seed"), &d.seed) ||
        !vec3_f32(t.get("pspeed"), d.ps) || !num_i64(t.get("overrides"), &ov) || !num_i64(t.get("lost"), &d.lost))
      return SWE_EINVAL;
    d.ov = (int32_t)ov;
    int32_t nx = 0, ny = 0;
    if ((r = swe_get_terrain(ctx, k, &nx, &ny, nullptr, nullptr, nullptr))) return r;
    if (nx != d.nx || ny != d.ny) return SWE_EINVAL;
  }
  if ((r = swe_set_params(ctx, &p))) return r;
  for (int32_t k = 0; k < nterr; k++) {
    const T& d = ts[(size_t)k];
    const size_t n = (size_t)d.nx * d.ny;
    std::vector<int64_t> h(n), s;
    std::vector<float> mf(n), mo(3 * n);
    bool ok = std::fread(h.data(), 8, n, f.f) == n;
    if (ok && p.layers == 2) {
      s.resize(n);
      ok = std::fread(s.data(), 8, n, f.f) == n;
    }
    if (ok) ok = std::fread(mf.data(), 4, n, f.f) == n && std::fread(mo.data(), 4, 3 * n, f.f) == 3 * n;
    if (!ok) return SWE_EINVAL;
    if (d.ov & 3) {
      if ((r = swe_set_terrain(ctx, k, (d.ov & 1) ? d.seed : p.seed, (d.ov & 2) ? d.ps : nullptr))) return r;
    }
    if ((r = swe_upload_terrain_master(ctx, k, h.data(), p.layers == 2 ? s.data() : nullptr))) return r;
    if ((r = swe_upload_terrain_fields(ctx, k, mf.data(), mo.data()))) return r;
    if ((r = swe_set_terrain_lost(ctx, k, d.lost))) return r;
  }
  return swe_set_cycle(ctx, cycle);
}
