// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// C-ABI wrappers around the parts of K-Litho that compile from their own
// source files with no external library: common/file_io.cpp,
// common/mask.cpp and K-Litho-TCC/source.cpp.  oracle/Makefile compiles those
// files where they lie under the reference tree together with this file into
// oracle/_ref/libklref.so.  The tests use it to pin the oracle's and the
// library's source generators, mask placement and kernel_info.txt handling
// against the reference's own code.  klitho_tcc.cpp / klitho_socs.cpp need
// FFTW3 and LAPACK headers that this image does not have, so they are not
// built.
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "file_io.hpp"
#include "mask.hpp"
#include "source.hpp"

extern "C" {

// source.cpp:105-142 (createSource), normalisation included
void ref_create_source(float* out, int srcSize, const char* srcType, float inner,
                       float outer, float radius, float offset, int onXAxis, float ptX,
                       float ptY, const char* file) {
  std::vector<float> src(srcSize * srcSize, 0.0f);
  Annular an{inner, outer};
  Dipole di{radius, offset, onXAxis != 0};
  CrossQuadrupole cq{radius, offset};
  Point pt{ptX, ptY};
  std::stringstream ss;
  createSource(src, srcSize, srcType, an, di, cq, pt, file ? file : "", ss);
  std::memcpy(out, src.data(), sizeof(float) * srcSize * srcSize);
}

// mask.cpp:31-60 (createMask)
void ref_create_mask(float* out, int Lx, int Ly, int msx, int msy, const char* type,
                     int lw, int sw, int horiz, const char* file, float dose) {
  std::vector<float> msk(Lx * Ly, 0.0f);
  LineSpace ls;
  ls.isHorizontal = horiz != 0;
  ls.lineWidth = lw;
  ls.spaceWidth = sw;
  std::stringstream ss;
  createMask(msk, Lx, Ly, msx, msy, type, ls, file ? file : "", ss, dose);
  std::memcpy(out, msk.data(), sizeof(float) * Lx * Ly);
}

// file_io.cpp:262-280 (writeKernelInfo)
void ref_write_kernel_info(const char* path, int sx, int sy, int nk, const float* scales) {
  std::vector<float> s(scales, scales + nk);
  writeKernelInfo(path, sx, sy, nk, s);
}

// file_io.cpp:282-307 (readKernelInfo); returns nk, fills up to cap scales
int ref_read_kernel_info(const char* path, int* sx, int* sy, float* scales, int cap) {
  int nk = 0;
  std::vector<float> s;
  readKernelInfo(path, *sx, *sy, nk, s);
  for (int i = 0; i < nk && i < cap && i < (int)s.size(); ++i) scales[i] = s[i];
  return nk;
}

// file_io.cpp:232-260 (readFloatArrayFromBinary); returns elements kept
int ref_read_float_bin(const char* path, float* out, int size) {
  std::vector<float> d;
  readFloatArrayFromBinary(path, d, size);
  std::memcpy(out, d.data(), sizeof(float) * size);
  return size;
}

// file_io.cpp:137-150 (writeFloatArrayToBinary)
void ref_write_float_bin(const char* path, const float* data, int size) {
  std::vector<float> d(data, data + size);
  writeFloatArrayToBinary(path, d, size);
}

// file_io.cpp:7-135: the three PNG writers of the tools
void ref_png_gray(const char* path, int sx, int sy, const float* data, float mn, float mx) {
  std::vector<float> d(data, data + (size_t)sx * sy);
  writeFloatArrayToPNGWithGrayscale(sx, sy, d, mn, mx, path);
}
void ref_png_7colors(const char* path, int sx, int sy, const float* data, float mn, float mx) {
  std::vector<float> d(data, data + (size_t)sx * sy);
  writeFloatArrayToPNGWith7Colors(sx, sy, d, path, mn, mx);
}
void ref_png_continuous(const char* path, int sx, int sy, const float* data, float mn,
                        float mx) {
  std::vector<float> d(data, data + (size_t)sx * sy);
  writeFloatArrayToPNGWithContinuousColor(sx, sy, d, path, mn, mx);
}

}  // extern "C"
