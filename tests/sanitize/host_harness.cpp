// host_harness.cpp -- TEST INFRASTRUCTURE: drives the host-only boundary code
// of libklt.so (sources, masks, sizing, file formats, PNG views) through
// ordinary and malformed inputs, for the AddressSanitizer / UndefinedBehavior
// build of tests/test_host_sanitize.py.  Exit status 0 when every call
// returned what it should; the sanitizers abort on any memory or UB error.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "klt.h"

static int fails = 0;
#define EXPECT(c)                                                  \
  do {                                                             \
    if (!(c)) {                                                    \
      std::fprintf(stderr, "expectation failed: %s (line %d)\n", #c, __LINE__); \
      ++fails;                                                     \
    }                                                              \
  } while (0)

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  // sources of every type and small / even / odd sizes, the point source out
  // of the grid, an unknown type
  const char* types[] = {"Annular", "Dipole", "CrossQuadrupole", "Point"};
  for (int n : {1, 2, 3, 4, 51, 101})
    for (const char* t : types) {
      std::vector<float> s((size_t)n * n, 0.f);
      EXPECT(klt_source_create(s.data(), n, t, 0.6f, 0.9f, 0.2f, 0.7f, 1, 0.6f, 0.2f, "") == KLT_OK);
      (void)klt_outer_sigma(s.data(), n);
    }
  {
    std::vector<float> s(51 * 51, 0.f);
    // a point outside the grid is refused (the reference writes out of bounds,
    // source.cpp:122-129); an unknown type leaves the zero source, as there
    EXPECT(klt_source_create(s.data(), 51, "Point", 0, 0, 0, 0, 0, 7.0f, -9.0f, "") != KLT_OK);
    EXPECT(klt_source_create(s.data(), 51, "NoSuchType", 0, 0, 0, 0, 0, 0, 0, "") == KLT_OK);
  }
  // masks: line/space at many sizes and periods, import of text and binary
  // files that are short, long, malformed or missing
  for (int L : {1, 7, 64, 65})
    for (int ms : {1, 5, L}) {
      if (ms > L) continue;
      std::vector<float> m((size_t)L * L, 0.f);
      EXPECT(klt_mask_create(m.data(), L, L, ms, ms, "LineSpace", 3, 2, L & 1, "", 1.0f) == KLT_OK);
    }
  const std::string txt = dir + "/m.txt", bin = dir + "/m.bin", bad = dir + "/bad.txt";
  {
    FILE* f = std::fopen(txt.c_str(), "w");
    for (int i = 0; i < 30; ++i) std::fprintf(f, "%g ", (float)(i % 3));
    std::fclose(f);
    f = std::fopen(bad.c_str(), "w");
    std::fprintf(f, "1 2 three 4");
    std::fclose(f);
    std::vector<float> v(50, 0.5f);
    EXPECT(klt_write_floats(bin.c_str(), v.data(), 50) == KLT_OK);
  }
  for (int ms : {4, 6, 8}) {  // 16 / 36 / 64 values wanted from 30 (text) or 50 (binary)
    std::vector<float> m(16 * 16, 0.f);
    EXPECT(klt_mask_create(m.data(), 16, 16, ms, ms, "Import", 0, 0, 0, txt.c_str(), 1.0f) == KLT_OK);
    EXPECT(klt_mask_create(m.data(), 16, 16, ms, ms, "Import", 0, 0, 0, bin.c_str(), 1.0f) == KLT_OK);
  }
  {
    std::vector<float> m(16 * 16, 0.f);
    EXPECT(klt_mask_create(m.data(), 16, 16, 4, 4, "Import", 0, 0, 0, bad.c_str(), 1.0f) != KLT_OK);
    EXPECT(klt_mask_create(m.data(), 16, 16, 4, 4, "Import", 0, 0, 0, (dir + "/none.bin").c_str(), 1.0f) != KLT_OK);
    std::vector<float> r(100);
    EXPECT(klt_read_floats((dir + "/none.txt").c_str(), r.data(), 100) != KLT_OK);
    EXPECT(klt_read_floats(bin.c_str(), r.data(), 100) == KLT_OK);  // short: zero-padded
    EXPECT(klt_read_floats(bin.c_str(), r.data(), 10) == KLT_OK);   // long: truncated
  }
  // sizing
  int nx = 0, ny = 0;
  EXPECT(klt_tcc_sizing(1.35f, 1088, 1088, 0.88f, 193.f, &nx, &ny) == KLT_OK && nx == 14);
  EXPECT(klt_tcc_sizing(1.35f, 0, 10, 0.5f, 193.f, &nx, &ny) != KLT_OK);
  // kernel_info: round trip, fewer eigenvalue lines than kernels, capacity
  // below the count, missing file
  {
    const std::string ki = dir + "/kernel_info.txt";
    std::vector<float> sc = {3.f, 2.f, 1.f};
    EXPECT(klt_write_kernel_info(ki.c_str(), 33, 33, 3, sc.data()) == KLT_OK);
    int sx = 0, sy = 0;
    float got[1];
    EXPECT(klt_read_kernel_info(ki.c_str(), &sx, &sy, got, 1) == 3 && sx == 33 && got[0] == 3.f);
    FILE* f = std::fopen(ki.c_str(), "w");
    std::fprintf(f, "- Kernel Size: 5x7\n- Number of Kernels: 4\nEigenvalue 0: 1\nEigenvalue 9: 2\n");
    std::fclose(f);
    float four[4];
    EXPECT(klt_read_kernel_info(ki.c_str(), &sx, &sy, four, 4) == 4 && sy == 7);
    EXPECT(klt_read_kernel_info((dir + "/no_info.txt").c_str(), &sx, &sy, four, 4) < 0);
  }
  // PNG views, including a one-pixel image, NaN and a flat range
  {
    std::vector<float> d = {0.f, NAN, INFINITY, -INFINITY, 0.1f, 0.2f};
    for (int k = 0; k < 3; ++k) {
      const std::string p = dir + "/v" + std::to_string(k) + ".png";
      auto fn = k == 0 ? klt_write_png_gray : k == 1 ? klt_write_png_7colors : klt_write_png_continuous;
      EXPECT(fn(p.c_str(), 3, 2, d.data(), 0.f, 0.225f) == KLT_OK);
      EXPECT(fn(p.c_str(), 1, 1, d.data(), 1.f, 1.f) == KLT_OK);
      EXPECT(fn((dir + "/no/dir/x.png").c_str(), 3, 2, d.data(), 0.f, 1.f) != KLT_OK);
    }
    std::vector<float> big(300 * 250, 0.3f);  // more than one stored deflate block
    EXPECT(klt_write_png_continuous((dir + "/big.png").c_str(), 300, 250, big.data(), 0.f, 1.f) == KLT_OK);
  }
  std::printf("host harness: %d failed expectations\n", fails);
  return fails ? 1 : 0;
}
