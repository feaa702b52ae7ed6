// swe_state.cpp -- checkpoint / resume of a context's exact state.
//
// The reference keeps no state beyond its seed (SimpleWindErosion.cpp:27-34).
// A checkpoint here is the exact int64 master layers plus everything the
// next cycle depends on, so that load + erode continues bit for bit:
//
//   offset  size          field
//   0       8             magic "SWESTAT1"
//   8       4             nx (int32, little endian)
//   12      4             ny
//   16      80            swe_params, raw
//   96      8             cycle counter (uint64)
//   104     8             lost mass (int64, fixed point 2^-32)
//   112     8*nx*ny       height master (int64, fixed point)
//   ...     8*nx*ny       sediment master (int64; present when layers == 2)
//   ...     4*4*nx*ny     EMA fields massflow, momentum x, y, z (float32)
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/swe.h"

namespace {
const char kMagic[8] = {'S', 'W', 'E', 'S', 'T', 'A', 'T', '1'};
static_assert(sizeof(swe_params) == 80, "swe_params layout");
}  // namespace

extern "C" int swe_save_state(swe_ctx* ctx, const char* path) {
  if (!ctx || !path) return SWE_EINVAL;
  swe_params p;
  int r;
  if ((r = swe_get_params(ctx, &p))) return r;
  swe_audit_t a;
  if ((r = swe_audit(ctx, &a))) return r;
  int32_t nx = 0, ny = 0;
  if ((r = swe_get_size(ctx, &nx, &ny))) return r;
  const size_t n = (size_t)nx * ny;
  std::vector<int64_t> h(n), s(n);
  std::vector<float> mf(n), mo(3 * n);
  if ((r = swe_download_master(ctx, h.data(), s.data()))) return r;
  if ((r = swe_download_layers(ctx, nullptr, nullptr, mf.data(), mo.data()))) return r;
  FILE* f = std::fopen(path, "wb");
  if (!f) return SWE_EINVAL;
  bool ok = std::fwrite(kMagic, 1, 8, f) == 8 && std::fwrite(&nx, 4, 1, f) == 1 && std::fwrite(&ny, 4, 1, f) == 1 &&
            std::fwrite(&p, sizeof(p), 1, f) == 1 && std::fwrite(&a.cycle, 8, 1, f) == 1 &&
            std::fwrite(&a.lost, 8, 1, f) == 1 && std::fwrite(h.data(), 8, n, f) == n;
  if (ok && p.layers == 2) ok = std::fwrite(s.data(), 8, n, f) == n;
  if (ok) ok = std::fwrite(mf.data(), 4, n, f) == n && std::fwrite(mo.data(), 4, 3 * n, f) == 3 * n;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? 0 : SWE_EINVAL;
}

extern "C" int swe_load_state(swe_ctx* ctx, const char* path) {
  if (!ctx || !path) return SWE_EINVAL;
  int32_t nx = 0, ny = 0;
  int r;
  if ((r = swe_get_size(ctx, &nx, &ny))) return r;
  FILE* f = std::fopen(path, "rb");
  if (!f) return SWE_EINVAL;
  char magic[8];
  int32_t fnx = 0, fny = 0;
  swe_params p;
  uint64_t cycle = 0;
  int64_t lost = 0;
  bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kMagic, 8) == 0 && std::fread(&fnx, 4, 1, f) == 1 &&
            std::fread(&fny, 4, 1, f) == 1 && fnx == nx && fny == ny && std::fread(&p, sizeof(p), 1, f) == 1 &&
            std::fread(&cycle, 8, 1, f) == 1 && std::fread(&lost, 8, 1, f) == 1;
  const size_t n = (size_t)nx * ny;
  std::vector<int64_t> h(n), s;
  std::vector<float> mf(n), mo(3 * n);
  if (ok) ok = std::fread(h.data(), 8, n, f) == n;
  if (ok && p.layers == 2) {
    s.resize(n);
    ok = std::fread(s.data(), 8, n, f) == n;
  }
  if (ok) ok = std::fread(mf.data(), 4, n, f) == n && std::fread(mo.data(), 4, 3 * n, f) == 3 * n;
  std::fclose(f);
  if (!ok) return SWE_EINVAL;
  if ((r = swe_set_params(ctx, &p))) return r;
  if ((r = swe_upload_master(ctx, h.data(), p.layers == 2 ? s.data() : nullptr))) return r;
  if ((r = swe_upload_fields(ctx, mf.data(), mo.data()))) return r;
  if ((r = swe_set_cycle(ctx, cycle))) return r;
  return swe_set_lost(ctx, lost);
}
