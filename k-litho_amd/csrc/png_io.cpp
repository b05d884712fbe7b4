// png_io.cpp -- the PNG views the reference's tools write (common/file_io.cpp:
// 7-135): grayscale, the 7-level partition of image_s.png and the continuous
// blue-to-red map.  The value-to-colour maps repeat the reference's float /
// double decisions (this file is compiled with -ffp-contract=off), so every
// pixel equals the reference's; the files are encoded here as 8-bit PNGs whose
// zlib stream uses stored (uncompressed) deflate blocks.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "klt_error.h"

namespace {

using klt::set_error;

uint32_t crc_table[256];
bool crc_ready = false;

void crc_init() {
  for (uint32_t n = 0; n < 256; ++n) {
    uint32_t c = n;
    for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
    crc_table[n] = c;
  }
  crc_ready = true;
}

uint32_t crc32(const unsigned char* p, size_t n, uint32_t c = 0xFFFFFFFFu) {
  for (size_t i = 0; i < n; ++i) c = crc_table[(c ^ p[i]) & 0xFF] ^ (c >> 8);
  return c;
}

void put_u32(std::vector<unsigned char>& v, uint32_t x) {
  v.push_back((unsigned char)(x >> 24));
  v.push_back((unsigned char)(x >> 16));
  v.push_back((unsigned char)(x >> 8));
  v.push_back((unsigned char)x);
}

void chunk(std::vector<unsigned char>& out, const char* type, const std::vector<unsigned char>& d) {
  put_u32(out, (uint32_t)d.size());
  const size_t start = out.size();
  out.insert(out.end(), type, type + 4);
  out.insert(out.end(), d.begin(), d.end());
  put_u32(out, crc32(out.data() + start, out.size() - start) ^ 0xFFFFFFFFu);
}

// 8-bit PNG of sy rows of sx pixels with `ch` channels (1 gray, 3 RGB)
int write_png(const char* path, int sx, int sy, int ch, const std::vector<unsigned char>& px) {
  if (!crc_ready) crc_init();
  std::vector<unsigned char> raw;  // filter byte 0 before every row
  raw.reserve((size_t)sy * (sx * ch + 1));
  for (int y = 0; y < sy; ++y) {
    raw.push_back(0);
    raw.insert(raw.end(), px.begin() + (size_t)y * sx * ch, px.begin() + (size_t)(y + 1) * sx * ch);
  }
  std::vector<unsigned char> z = {0x78, 0x01};
  uint32_t a = 1, b = 0;  // Adler-32
  for (unsigned char c : raw) {
    a = (a + c) % 65521u;
    b = (b + a) % 65521u;
  }
  for (size_t off = 0; off < raw.size();) {  // (raw holds at least one filter byte)
    const size_t n = std::min<size_t>(65535, raw.size() - off);
    const bool last = off + n >= raw.size();
    z.push_back(last ? 1 : 0);
    z.push_back((unsigned char)(n & 0xFF));
    z.push_back((unsigned char)(n >> 8));
    z.push_back((unsigned char)(~n & 0xFF));
    z.push_back((unsigned char)((~n >> 8) & 0xFF));
    z.insert(z.end(), raw.begin() + off, raw.begin() + off + n);
    off += n;
  }
  put_u32(z, (b << 16) | a);
  std::vector<unsigned char> ihdr;
  put_u32(ihdr, (uint32_t)sx);
  put_u32(ihdr, (uint32_t)sy);
  ihdr.push_back(8);
  ihdr.push_back(ch == 1 ? 0 : 2);
  ihdr.push_back(0);
  ihdr.push_back(0);
  ihdr.push_back(0);
  std::vector<unsigned char> out = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
  chunk(out, "IHDR", ihdr);
  chunk(out, "IDAT", z);
  chunk(out, "IEND", {});
  FILE* f = std::fopen(path, "wb");
  bool ok = f && std::fwrite(out.data(), 1, out.size(), f) == out.size();
  if (f && std::fclose(f) != 0) ok = false;
  if (!ok) return set_error(KLT_ERR_IO, std::string("Error: Failed to write PNG: ") + path);
  return KLT_OK;
}

// std::max(0.0f, std::min(1.0f, q)) with the standard library's comparisons
// (a NaN q gives 1, like the reference)
inline float clamp01(float q) {
  const float t = (q < 1.0f) ? q : 1.0f;
  return (0.0f < t) ? t : 0.0f;
}

int check_args(const char* path, int sx, int sy, const float* data, const char* who) {
  if (!path || sx <= 0 || sy <= 0 || !data) return set_error(KLT_ERR_ARG, std::string(who) + ": bad argument");
  return KLT_OK;
}

}  // namespace

extern "C" {

// file_io.cpp:7-23 (writeFloatArrayToPNGWithGrayscale)
int klt_write_png_gray(const char* path, int sx, int sy, const float* data, float min,
                       float max) {
  KLT_TRY(check_args(path, sx, sy, data, "klt_write_png_gray"));
  const size_t n = (size_t)sx * sy;
  std::vector<unsigned char> px(n);
  for (size_t i = 0; i < n; ++i) {
    const float v = clamp01((data[i] - min) / (max - min));
    px[i] = static_cast<unsigned char>(v * 255.0f);
  }
  return write_png(path, sx, sy, 1, px);
}

// file_io.cpp:25-80 (writeFloatArrayToPNGWith7Colors)
int klt_write_png_7colors(const char* path, int sx, int sy, const float* data, float min,
                          float max) {
  KLT_TRY(check_args(path, sx, sy, data, "klt_write_png_7colors"));
  static const unsigned char rgb[7][3] = {{0, 0, 255},   {0, 127, 255}, {0, 255, 127},
                                          {0, 255, 0},   {127, 255, 0}, {255, 127, 0},
                                          {255, 0, 0}};
  const size_t n = (size_t)sx * sy;
  std::vector<unsigned char> px(3 * n);
  for (size_t i = 0; i < n; ++i) {
    int lev = 3;  // min == max: uniform green
    if (min != max) {
      const double v = clamp01((data[i] - min) / (max - min));
      lev = v < 1.0 / 6.0 ? 0 : v < 2.0 / 6.0 ? 1 : v < 3.0 / 6.0 ? 2 : v < 4.0 / 6.0 ? 3
          : v < 5.0 / 6.0 ? 4 : v < 6.0 / 6.0 ? 5 : 6;
    }
    for (int c = 0; c < 3; ++c) px[3 * i + c] = rgb[lev][c];
  }
  return write_png(path, sx, sy, 3, px);
}

// file_io.cpp:82-135 (writeFloatArrayToPNGWithContinuousColor)
int klt_write_png_continuous(const char* path, int sx, int sy, const float* data, float min,
                             float max) {
  KLT_TRY(check_args(path, sx, sy, data, "klt_write_png_continuous"));
  const size_t n = (size_t)sx * sy;
  std::vector<unsigned char> px(3 * n);
  for (size_t i = 0; i < n; ++i) {
    unsigned char R = 0, G = 255, B = 0;  // min == max: uniform green
    if (min != max) {
      const double v = clamp01((data[i] - min) / (max - min));
      R = G = B = 0;
      if (v < 1.0 / 6.0) {  // blue -> aqua
        B = 255;
        G = static_cast<unsigned char>(v * 6 * 127);
      } else if (v < 2.0 / 6.0) {  // aqua -> bright green
        G = 127 + static_cast<unsigned char>((v - 1.0 / 6.0) * 6 * 128);
        B = static_cast<unsigned char>(255 - (v - 1.0 / 6.0) * 6 * 127);
      } else if (v < 3.0 / 6.0) {  // bright green -> green
        G = 255;
        B = static_cast<unsigned char>(128 - (v - 2.0 / 6.0) * 6 * 128);
      } else if (v < 4.0 / 6.0) {  // green -> yellow
        R = static_cast<unsigned char>((v - 3.0 / 6.0) * 6 * 127);
        G = 255;
      } else if (v < 5.0 / 6.0) {  // yellow -> orange
        R = 127 + static_cast<unsigned char>((v - 4.0 / 6.0) * 6 * 128);
        G = static_cast<unsigned char>(255 - (v - 4.0 / 6.0) * 6 * 127);
      } else {  // orange -> red
        R = 255;
        G = static_cast<unsigned char>(128 - (v - 5.0 / 6.0) * 6 * 128);
      }
    }
    px[3 * i] = R;
    px[3 * i + 1] = G;
    px[3 * i + 2] = B;
  }
  return write_png(path, sx, sy, 3, px);
}

}  // extern "C"
