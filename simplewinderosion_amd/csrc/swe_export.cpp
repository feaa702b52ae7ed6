// swe_export.cpp -- flow-field and map products written to files.
//
// The reference draws its wind map into a texture every frame
// (SimpleWindErosion.cpp:224-230) and keeps the EMA flow fields in its cells
// (world.hpp:69-76); it writes no files.  Two formats, both written without
// compression so that they need no library:
//
//   SWE_EXPORT_NPZ: a numpy .npz archive (a ZIP of .npy files, method
//     "stored"): height, sediment, massflow (nx, ny) float32; momentum and
//     normals (3, nx, ny) float32; windmap (nx, ny, 4) float32; height_fx and
//     sediment_fx (nx, ny) int64, the exact fixed-point (2^-32) masters.
//   SWE_EXPORT_PNG: the wind map as an 8-bit RGBA PNG, nx pixels wide and ny
//     rows high, pixel (x, y) = cell (x, y); channel = round(255 v), v
//     clamped to [0, 1].
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/swe.h"

namespace {

// CRC-32 (ISO 3309 / ITU-T V.42 polynomial 0xEDB88320, reflected), as ZIP and
// PNG define it
uint32_t crc32_update(uint32_t crc, const uint8_t* p, size_t n) {
  static uint32_t table[256];
  static bool init = false;
  if (!init) {
    for (uint32_t i = 0; i < 256; i++) {
      uint32_t c = i;
      for (int k = 0; k < 8; k++) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
      table[i] = c;
    }
    init = true;
  }
  crc = ~crc;
  for (size_t i = 0; i < n; i++) crc = table[(crc ^ p[i]) & 0xffu] ^ (crc >> 8);
  return ~crc;
}

void put16(std::vector<uint8_t>& b, uint32_t v) {
  b.push_back((uint8_t)(v & 0xff));
  b.push_back((uint8_t)((v >> 8) & 0xff));
}
void put32(std::vector<uint8_t>& b, uint32_t v) {
  put16(b, v & 0xffffu);
  put16(b, v >> 16);
}
void put32be(std::vector<uint8_t>& b, uint32_t v) {
  b.push_back((uint8_t)(v >> 24));
  b.push_back((uint8_t)((v >> 16) & 0xff));
  b.push_back((uint8_t)((v >> 8) & 0xff));
  b.push_back((uint8_t)(v & 0xff));
}

// .npy version 1.0: magic, header length, header dict padded with spaces to
// a multiple of 64 bytes ending in a newline, then the raw little-endian data
std::vector<uint8_t> npy(const char* descr, const std::vector<int64_t>& shape, const void* data, size_t bytes) {
  std::string dict = std::string("{'descr': '") + descr + "', 'fortran_order': False, 'shape': (";
  for (size_t i = 0; i < shape.size(); i++) dict += std::to_string(shape[i]) + (shape.size() == 1 || i + 1 < shape.size() ? ", " : "");
  dict += "), }";
  const size_t pre = 10;
  size_t total = pre + dict.size() + 1;
  const size_t pad = (64 - total % 64) % 64;
  dict += std::string(pad, ' ') + "\n";
  std::vector<uint8_t> b;
  const uint8_t magic[6] = {0x93, 'N', 'U', 'M', 'P', 'Y'};
  b.insert(b.end(), magic, magic + 6);
  b.push_back(1);
  b.push_back(0);
  put16(b, (uint32_t)dict.size());
  b.insert(b.end(), dict.begin(), dict.end());
  const uint8_t* p = (const uint8_t*)data;
  b.insert(b.end(), p, p + bytes);
  return b;
}

struct ZipEntry {
  std::string name;
  uint32_t crc, size, offset;
};

// ZIP archive of stored (uncompressed) members, each below 4 GiB
class ZipWriter {
 public:
  explicit ZipWriter(FILE* f) : f_(f) {}
  bool add(const std::string& name, const std::vector<uint8_t>& data) {
    if (data.size() >= 0xffffffffu || off_ + data.size() + 64 + name.size() >= 0xffffffffull) return false;
    ZipEntry e{name, crc32_update(0, data.data(), data.size()), (uint32_t)data.size(), (uint32_t)off_};
    std::vector<uint8_t> h;
    put32(h, 0x04034b50u);  // local file header
    put16(h, 20);           // version needed
    put16(h, 0);            // flags
    put16(h, 0);            // method: stored
    put16(h, 0);            // time
    put16(h, 0x21);         // date: 1980-01-01
    put32(h, e.crc);
    put32(h, e.size);
    put32(h, e.size);
    put16(h, (uint32_t)name.size());
    put16(h, 0);
    h.insert(h.end(), name.begin(), name.end());
    if (!write(h) || !write(data)) return false;
    entries_.push_back(e);
    return true;
  }
  bool finish() {
    std::vector<uint8_t> c;
    for (const auto& e : entries_) {
      put32(c, 0x02014b50u);  // central directory header
      put16(c, 20);
      put16(c, 20);
      put16(c, 0);
      put16(c, 0);
      put16(c, 0);
      put16(c, 0x21);
      put32(c, e.crc);
      put32(c, e.size);
      put32(c, e.size);
      put16(c, (uint32_t)e.name.size());
      put16(c, 0);
      put16(c, 0);
      put16(c, 0);
      put16(c, 0);
      put32(c, 0);
      put32(c, e.offset);
      c.insert(c.end(), e.name.begin(), e.name.end());
    }
    const uint64_t cd_off = off_;
    if (!write(c)) return false;
    std::vector<uint8_t> z;
    put32(z, 0x06054b50u);  // end of central directory
    put16(z, 0);
    put16(z, 0);
    put16(z, (uint32_t)entries_.size());
    put16(z, (uint32_t)entries_.size());
    put32(z, (uint32_t)c.size());
    put32(z, (uint32_t)cd_off);
    put16(z, 0);
    return write(z);
  }

 private:
  bool write(const std::vector<uint8_t>& b) {
    if (std::fwrite(b.data(), 1, b.size(), f_) != b.size()) return false;
    off_ += b.size();
    return true;
  }
  FILE* f_;
  uint64_t off_ = 0;
  std::vector<ZipEntry> entries_;
};

// PNG chunk: length, type, data, CRC over type and data
void png_chunk(std::vector<uint8_t>& out, const char* type, const std::vector<uint8_t>& data) {
  put32be(out, (uint32_t)data.size());
  std::vector<uint8_t> td(type, type + 4);
  td.insert(td.end(), data.begin(), data.end());
  out.insert(out.end(), td.begin(), td.end());
  put32be(out, crc32_update(0, td.data(), td.size()));
}

// zlib stream of stored deflate blocks (RFC 1950 / 1951), Adler-32 trailer
std::vector<uint8_t> zlib_stored(const std::vector<uint8_t>& raw) {
  std::vector<uint8_t> z = {0x78, 0x01};
  size_t pos = 0;
  do {
    const size_t n = raw.size() - pos < 65535 ? raw.size() - pos : 65535;
    const bool last = pos + n == raw.size();
    z.push_back(last ? 1 : 0);
    put16(z, (uint32_t)n);
    put16(z, (uint32_t)(~n & 0xffffu));
    z.insert(z.end(), raw.begin() + (long)pos, raw.begin() + (long)(pos + n));
    pos += n;
  } while (pos < raw.size());
  uint32_t a = 1, b = 0;
  for (uint8_t v : raw) {
    a = (a + v) % 65521u;
    b = (b + a) % 65521u;
  }
  put32be(z, (b << 16) | a);
  return z;
}

uint8_t to_u8(float v) {
  if (!(v > 0.0f)) return 0;
  if (v >= 1.0f) return 255;
  return (uint8_t)std::lround(255.0 * (double)v);
}

int export_png(swe_ctx* ctx, int32_t k, int32_t nx, int32_t ny, FILE* f) {
  const size_t n = (size_t)nx * ny;
  std::vector<float> wm(4 * n);
  int r;
  if ((r = swe_windmap_terrain(ctx, k, wm.data()))) return r;
  // rows of the image run along x: row y holds cells (0..nx-1, y)
  std::vector<uint8_t> raw;
  raw.reserve((size_t)ny * (1 + 4 * (size_t)nx));
  for (int32_t y = 0; y < ny; y++) {
    raw.push_back(0);  // filter: none
    for (int32_t x = 0; x < nx; x++) {
      const float* c = &wm[4 * ((size_t)x * ny + y)];
      for (int q = 0; q < 4; q++) raw.push_back(to_u8(c[q]));
    }
  }
  std::vector<uint8_t> png = {0x89, 'P', 'N', 'G', 0x0d, 0x0a, 0x1a, 0x0a};
  std::vector<uint8_t> ihdr;
  put32be(ihdr, (uint32_t)nx);
  put32be(ihdr, (uint32_t)ny);
  ihdr.push_back(8);  // bit depth
  ihdr.push_back(6);  // colour type: RGBA
  ihdr.push_back(0);
  ihdr.push_back(0);
  ihdr.push_back(0);
  png_chunk(png, "IHDR", ihdr);
  png_chunk(png, "IDAT", zlib_stored(raw));
  png_chunk(png, "IEND", {});
  return std::fwrite(png.data(), 1, png.size(), f) == png.size() ? 0 : SWE_EINVAL;
}

int export_npz(swe_ctx* ctx, int32_t k, int32_t nx, int32_t ny, FILE* f) {
  const size_t n = (size_t)nx * ny;
  std::vector<float> h(n), s(n), mf(n), mo(3 * n), wm(4 * n), nm(3 * n);
  std::vector<int64_t> hf(n), sf(n);
  int r;
  if ((r = swe_download_terrain(ctx, k, h.data(), s.data(), mf.data(), mo.data())) ||
      (r = swe_download_terrain_master(ctx, k, hf.data(), sf.data())) ||
      (r = swe_windmap_terrain(ctx, k, wm.data())) || (r = swe_normals_terrain(ctx, k, nm.data())))
    return r;
  const std::vector<int64_t> s2 = {nx, ny}, s3 = {3, nx, ny}, s4 = {nx, ny, 4};
  ZipWriter z(f);
  const bool ok = z.add("height.npy", npy("<f4", s2, h.data(), 4 * n)) &&
                  z.add("sediment.npy", npy("<f4", s2, s.data(), 4 * n)) &&
                  z.add("massflow.npy", npy("<f4", s2, mf.data(), 4 * n)) &&
                  z.add("momentum.npy", npy("<f4", s3, mo.data(), 12 * n)) &&
                  z.add("windmap.npy", npy("<f4", s4, wm.data(), 16 * n)) &&
                  z.add("normals.npy", npy("<f4", s3, nm.data(), 12 * n)) &&
                  z.add("height_fx.npy", npy("<i8", s2, hf.data(), 8 * n)) &&
                  z.add("sediment_fx.npy", npy("<i8", s2, sf.data(), 8 * n)) && z.finish();
  return ok ? 0 : SWE_EINVAL;
}

}  // namespace

extern "C" int swe_export_terrain(swe_ctx* ctx, int32_t k, const char* path, int32_t format) {
  if (!ctx || !path || (format != SWE_EXPORT_NPZ && format != SWE_EXPORT_PNG)) return SWE_EINVAL;
  int32_t nx = 0, ny = 0;
  int r;
  if ((r = swe_get_terrain(ctx, k, &nx, &ny, nullptr, nullptr, nullptr))) return r;
  FILE* f = std::fopen(path, "wb");
  if (!f) return SWE_EINVAL;
  r = format == SWE_EXPORT_NPZ ? export_npz(ctx, k, nx, ny, f) : export_png(ctx, k, nx, ny, f);
  const bool closed = std::fclose(f) == 0;
  if (r == 0 && !closed) r = SWE_EINVAL;
  if (r != 0) std::remove(path);
  return r;
}

extern "C" int swe_export(swe_ctx* ctx, const char* path, int32_t format) {
  int32_t nterr = 0;
  if (!ctx || swe_get_terrains(ctx, &nterr) || nterr != 1) return SWE_EINVAL;
  return swe_export_terrain(ctx, 0, path, format);
}
