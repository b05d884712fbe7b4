// host_io.cpp -- the host side of the drop-in boundary: source and mask
// generation, sizing, and the reference's file formats.  These are discrete
// float / integer computations whose results must be identical to the
// reference's (a flipped source cell or pupil sample changes the TCC), so
// they repeat the reference's operation order in float and this file is
// compiled with -ffp-contract=off.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "klt_error.h"

namespace {

using klt::set_error;

// K-Litho-TCC/source.cpp:5-24
void gen_annular(float* m, int n, float inner, float outer) {
  const int half = (n - 1) / 2;
  const int ri = (int)(inner * (float)half);
  const int ro = (int)(outer * (float)half);
  for (int y = 0; y < n; ++y)
    for (int x = 0; x < n; ++x) {
      const double d = std::sqrt(std::pow((double)(x - half), 2.0) +
                                 std::pow((double)(y - half), 2.0));
      m[y * n + x] = (d >= ri && d <= ro) ? 1.0f : 0.0f;
    }
}

// Discs of radius `rad` (cells) around `nc` centres: source.cpp:25-88
void gen_discs(float* m, int n, const int (*cen)[2], int nc, int rad) {
  for (int y = 0; y < n; ++y)
    for (int x = 0; x < n; ++x) {
      bool in = false;
      for (int i = 0; i < nc && !in; ++i) {
        const double dx = x - cen[i][0], dy = y - cen[i][1];
        in = std::sqrt(dx * dx + dy * dy) <= rad;
      }
      m[y * n + x] = in ? 1.0f : 0.0f;
    }
}

int read_bin(const std::string& path, float* out, int size) {
  // file_io.cpp:232-260: keeps min(size, bytes/4) values, zero-fills the rest
  std::ifstream f(path, std::ios::binary);
  if (!f.is_open()) return set_error(KLT_ERR_IO, "Error: Failed to open file " + path);
  f.seekg(0, std::ios::end);
  const long long bytes = (long long)f.tellg();
  f.seekg(0, std::ios::beg);
  std::vector<char> buf((size_t)std::max(0LL, bytes));
  if (bytes > 0 && !f.read(buf.data(), bytes))
    return set_error(KLT_ERR_IO, "Error: Failed to read data from file " + path);
  if (bytes != (long long)size * 4)
    std::cerr << "Warning: File size (" << bytes
              << " bytes) does not match the expected data size (" << size * sizeof(float)
              << " bytes)" << std::endl;
  const long long keep = std::min<long long>(size, bytes / 4);
  std::memset(out, 0, sizeof(float) * size);
  if (keep > 0) std::memcpy(out, buf.data(), sizeof(float) * keep);
  return KLT_OK;
}

int read_txt(const std::string& path, float* out, int size) {
  // file_io.cpp:204-230: whitespace-separated floats, resized to `size`
  std::ifstream f(path);
  if (!f.is_open()) return set_error(KLT_ERR_IO, "Error: Failed to open file " + path);
  std::vector<float> d;
  float v;
  while (f >> v) d.push_back(v);
  if (!f.eof()) return set_error(KLT_ERR_IO, "Error: Failed to read data from file " + path);
  if ((int)d.size() != size)
    std::cerr << "Warning: Data size (" << d.size() << ") does not match the expected size ("
              << size << ")" << std::endl;
  d.resize(size, 0.0f);
  std::memcpy(out, d.data(), sizeof(float) * size);
  return KLT_OK;
}

// The reference picks the reader by the text after the last '.' (source.cpp:132,
// mask.cpp:45): "bin" -> raw float32, anything else -> text.
int read_any(const std::string& path, float* out, int size) {
  const size_t dot = path.find_last_of('.');
  const std::string ext = path.substr(dot == std::string::npos ? 0 : dot + 1);
  return ext == "bin" ? read_bin(path, out, size) : read_txt(path, out, size);
}

}  // namespace

extern "C" {

// source.cpp:105-142 (createSource): generate or import, then normalise to
// sum 1 with a sequential float sum (source.cpp:89-103).  Unknown types give
// an all-zero source, as in the reference.
int klt_source_create(float* src, int srcSize, const char* srcType, float innerR,
                      float outerR, float radius, float offset, int onXAxis, float ptX,
                      float ptY, const char* sourceFile) {
  if (!src || !srcType || srcSize <= 0)
    return set_error(KLT_ERR_ARG, "klt_source_create: bad argument");
  const int n = srcSize;
  std::memset(src, 0, sizeof(float) * n * n);
  const std::string t(srcType);
  const int half = (n - 1) / 2;
  if (t == "Annular") {
    gen_annular(src, n, innerR, outerR);
  } else if (t == "Dipole") {
    const int off = (int)(offset * (float)half), rad = (int)(radius * (float)half);
    const int c = n / 2;  // source.cpp:61-62
    const int cen[2][2] = {{c + (onXAxis ? off : 0), c + (onXAxis ? 0 : off)},
                           {c - (onXAxis ? off : 0), c - (onXAxis ? 0 : off)}};
    gen_discs(src, n, cen, 2, rad);
  } else if (t == "CrossQuadrupole") {
    const int off = (int)(offset * (float)half), rad = (int)(radius * (float)half);
    const int cen[4][2] = {{half, half + off}, {half, half - off}, {half + off, half},
                           {half - off, half}};
    gen_discs(src, n, cen, 4, rad);
  } else if (t == "Point") {
    // source.cpp:124-128
    const float fx = ptX * (float)n, fy = (-1.0f * ptY) * (float)n;
    int ix = (int)std::round(fx / 2.0 + n / 2.0);
    int iy = (int)std::round(fy / 2.0 + n / 2.0);
    if (ptX >= 0) ix -= 1;
    if (-1.0f * ptY >= 0) iy -= 1;
    if (ix < 0 || ix >= n || iy < 0 || iy >= n)
      return set_error(KLT_ERR_ARG, "Point source outside the source grid");
    src[iy * n + ix] = 1.0f;
  } else if (t == "Import") {
    KLT_TRY(read_any(sourceFile ? sourceFile : "", src, n * n));
  }
  float sum = 0.0f;
  for (int i = 0; i < n * n; ++i) sum += src[i];
  if (sum != 0.0f)
    for (int i = 0; i < n * n; ++i) src[i] /= sum;
  return KLT_OK;
}

// mask.cpp:31-60 (createMask): LineSpace or Import, centred in Lx x Ly at
// offsets ((L - maskSize) / 2), times dose.  msk is overwritten.
int klt_mask_create(float* msk, int Lx, int Ly, int maskSizeX, int maskSizeY,
                    const char* maskType, int lineWidth, int spaceWidth, int isHorizontal,
                    const char* maskFile, float dose) {
  if (!msk || !maskType || Lx <= 0 || Ly <= 0 || maskSizeX <= 0 || maskSizeY <= 0 ||
      maskSizeX > Lx || maskSizeY > Ly)
    return set_error(KLT_ERR_ARG, "Error: Invalid input parameters. Please check the values.");
  std::vector<float> in((size_t)maskSizeX * maskSizeY, 0.0f);
  const std::string t(maskType);
  if (t == "LineSpace") {
    const int period = lineWidth + spaceWidth;
    if (period <= 0) return set_error(KLT_ERR_ARG, "LineSpace: lineWidth + spaceWidth <= 0");
    for (int y = 0; y < maskSizeY; ++y)
      for (int x = 0; x < maskSizeX; ++x) {
        const int r = isHorizontal ? (y % period) : (x % period);
        in[(size_t)y * maskSizeX + x] = r < lineWidth ? 1.0f : 0.0f;
      }
  } else if (t == "Import") {
    KLT_TRY(read_any(maskFile ? maskFile : "", in.data(), maskSizeX * maskSizeY));
  }
  std::memset(msk, 0, sizeof(float) * (size_t)Lx * Ly);
  const int dy = (Ly - maskSizeY) / 2, dx = (Lx - maskSizeX) / 2;
  for (int y = 0; y < maskSizeY; ++y)
    for (int x = 0; x < maskSizeX; ++x)
      msk[(size_t)(y + dy) * Lx + x + dx] = in[(size_t)y * maskSizeX + x] * dose;
  return KLT_OK;
}

// klitho_tcc.cpp:52-67 (getOuterSigma), unclamped
float klt_outer_sigma(const float* src, int srcSize) {
  const int c = (srcSize - 1) / 2;
  float maxd = 0.0f;
  for (int y = 0; y < srcSize; ++y)
    for (int x = 0; x < srcSize; ++x)
      if (src[y * srcSize + x] != 0.0f) {
        const float d = (float)std::sqrt((double)((y - c) * (y - c) + (x - c) * (x - c)));
        if (d > maxd) maxd = d;
      }
  return c > 0 ? maxd / (float)c : 0.0f;
}

// klitho_tcc.cpp:479-486: sigma clamped to 1, N = floor(f32(NA)*L*(1+sigma)/lambda)
int klt_tcc_sizing(float NA, int Lx, int Ly, float outerSigma, float lambda, int* Nx,
                   int* Ny) {
  if (!Nx || !Ny || Lx <= 0 || Ly <= 0 || lambda <= 0)
    return set_error(KLT_ERR_ARG, "klt_tcc_sizing: bad argument");
  if (outerSigma > 1.0f) outerSigma = 1.0f;
  const float s1 = 1.0f + outerSigma;
  float ax = NA * (float)Lx;
  ax = ax * s1;
  ax = ax / lambda;
  float ay = NA * (float)Ly;
  ay = ay * s1;
  ay = ay / lambda;
  *Nx = (int)std::floor(ax);
  *Ny = (int)std::floor(ay);
  return KLT_OK;
}

// file_io.cpp:232-260 / 204-230, chosen by extension like the reference
int klt_read_floats(const char* path, float* out, int size) {
  if (!path || !out || size < 0) return set_error(KLT_ERR_ARG, "klt_read_floats: bad argument");
  return read_any(path, out, size);
}

// file_io.cpp:137-150 (writeFloatArrayToBinary): raw little-endian float32
int klt_write_floats(const char* path, const float* data, int size) {
  if (!path || (!data && size > 0)) return set_error(KLT_ERR_ARG, "klt_write_floats: bad argument");
  std::ofstream f(path, std::ios::binary);
  if (!f.is_open()) return set_error(KLT_ERR_IO, std::string("Error: Failed to open file ") + path);
  f.write(reinterpret_cast<const char*>(data), (std::streamsize)size * sizeof(float));
  if (!f) return set_error(KLT_ERR_IO, std::string("Error: Failed to write data to file ") + path);
  return KLT_OK;
}

// file_io.cpp:262-280 (writeKernelInfo): same text, eigenvalues printed with
// the default 6 significant digits of an ostream.
int klt_write_kernel_info(const char* path, int sizeX, int sizeY, int nk, const float* scales) {
  if (!path || (nk > 0 && !scales)) return set_error(KLT_ERR_ARG, "klt_write_kernel_info: bad argument");
  std::ofstream f(path);
  if (!f) return set_error(KLT_ERR_IO, std::string("Error opening file for writing: ") + path);
  f << "# Kernel Information\n";
  f << "## Kernel Size and Count\n";
  f << "- Kernel Size: " << sizeX << "x" << sizeY << "\n";
  f << "- Number of Kernels: " << nk << "\n";
  f << "## Eigenvalues of Each Kernel\n";
  for (int i = 0; i < nk; ++i) f << "Eigenvalue " << i << ": " << scales[i] << "\n";
  return f ? KLT_OK : set_error(KLT_ERR_IO, std::string("write failed: ") + path);
}

// file_io.cpp:282-307 (readKernelInfo).  Returns the kernel count (>= 0) or an
// error; fills at most `cap` eigenvalues.
int klt_read_kernel_info(const char* path, int* sizeX, int* sizeY, float* scales, int cap) {
  if (!path || !sizeX || !sizeY) return set_error(KLT_ERR_ARG, "klt_read_kernel_info: bad argument");
  std::ifstream f(path);
  if (!f) return set_error(KLT_ERR_IO, std::string("Error opening file for reading: ") + path);
  int nk = 0;
  *sizeX = *sizeY = 0;
  std::vector<float> sc;
  std::string line;
  while (std::getline(f, line)) {
    if (line.find("Kernel Size:") != std::string::npos) {
      std::sscanf(line.c_str(), "- Kernel Size: %dx%d", sizeX, sizeY);
    } else if (line.find("Number of Kernels:") != std::string::npos) {
      std::sscanf(line.c_str(), "- Number of Kernels: %d", &nk);
    } else if (line.find("Eigenvalue") != std::string::npos) {
      int idx = -1;
      double v = 0.0;
      std::sscanf(line.c_str(), "Eigenvalue %d: %lf", &idx, &v);
      if (idx >= 0 && idx < nk) {
        sc.resize(nk, 0.0f);
        sc[idx] = (float)v;
      }
    }
  }
  for (int i = 0; i < nk && i < cap && scales; ++i) scales[i] = i < (int)sc.size() ? sc[i] : 0.0f;
  return nk;
}

}  // extern "C"
