// klitho_socs -- drop-in for K-Litho's SOCS tool (K-Litho-SOCS/klitho_socs.cpp:243-368)
// running the SOCS image on the GPU through libklt.
//
// Same 10 positional arguments, messages and exit codes; reads
// <krnDir>kernel_info.txt and <krnDir>krn_<i>_{r,i}.bin and writes
// ./out/image.bin (float32 Ly x Lx) and the reference's PNG views
// ./out/{mask,image,image_s}.png (file_io.cpp:7-135).  Environment:
//   KLT_SOCS_PATH  "fused" (default: split-precision matrix-core passes,
//                  N <= 15) or "plain" (FP32 single-image path).
//   KLT_DEVICE     GPU index (default 0).
#include <chrono>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "klt.h"

using namespace std;

namespace {

[[noreturn]] void die(const string& msg) {
  cerr << msg << endl;
  exit(1);
}

void check(int rc) {
  if (rc != KLT_OK) die(klt_last_error());
}

// getMax / getMin (klitho_socs.cpp:16-35)
float max_of(const vector<float>& v) {
  float m = -1e+30f;
  for (float x : v)
    if (m < x) m = x;
  return m;
}

float min_of(const vector<float>& v) {
  float m = 1e+30f;
  for (float x : v)
    if (m > x) m = x;
  return m;
}

}  // namespace

int main(int argc, char* argv[]) {
  cout << "************************************" << endl;
  cout << "* Lithography Simulation initiated..." << endl;
  cout << "************************************" << endl;
  auto start = chrono::system_clock::now();
  if (argc != 11) {
    cerr << "Error: Invalid number of arguments: " << argc << ". Expected 11 arguments." << endl;
    return 1;
  }
  const int Lx = stoi(argv[1]), Ly = stoi(argv[2]);
  const int maskSizeX = stoi(argv[3]), maskSizeY = stoi(argv[4]);
  const string maskType = argv[5];
  const int lineWidth = stoi(argv[6]), spaceWidth = stoi(argv[7]);
  const int isHorizontal = stoi(argv[8]);
  const string maskInFile = argv[9];
  const string KrnDir = argv[10];
  const float dose = 1.0f;
  if (Lx <= 0 || Ly <= 0 || maskSizeX <= 0 || maskSizeY <= 0 || Lx < maskSizeX ||
      Ly < maskSizeY) {
    cerr << "Error: Invalid input parameters. Please check the values." << endl;
    return 1;
  }
  int krnSizeX = 0, krnSizeY = 0;
  vector<float> scales(1 << 16);
  const int nk = klt_read_kernel_info((KrnDir + "kernel_info.txt").c_str(), &krnSizeX, &krnSizeY,
                                      scales.data(), (int)scales.size());
  if (nk < 0) die(klt_last_error());
  scales.resize(nk);
  if (krnSizeX % 2 == 0 || krnSizeY % 2 == 0) {
    cerr << "Error: Kernel size must be odd." << endl;
    return 1;
  }
  const int Nx = (krnSizeX - 1) / 2, Ny = (krnSizeY - 1) / 2;
  const int T = krnSizeX * krnSizeY;
  vector<float> kr((size_t)nk * T), ki((size_t)nk * T);
  for (int i = 0; i < nk; ++i) {
    const string fr = KrnDir + "krn_" + to_string(i) + "_r.bin";
    const string fi = KrnDir + "krn_" + to_string(i) + "_i.bin";
    cout << "Read Kernel " << i << " : " << fr << endl;
    check(klt_read_floats(fr.c_str(), kr.data() + (size_t)i * T, T));
    cout << "Read Kernel " << i << " : " << fi << endl;
    check(klt_read_floats(fi.c_str(), ki.data() + (size_t)i * T, T));
  }
  vector<float> msk((size_t)Lx * Ly, 0.0f);
  if (maskType == "Import") cout << "Read Mask: " << maskInFile << endl;
  check(klt_mask_create(msk.data(), Lx, Ly, maskSizeX, maskSizeY, maskType.c_str(), lineWidth,
                        spaceWidth, isHorizontal, maskInFile.c_str(), dose));
  check(klt_write_png_gray("./out/mask.png", Lx, Ly, msk.data(), min_of(msk), max_of(msk)));
  cout << endl << "Simulation Parameters:" << endl;
  cout << "-------------------" << endl;
  cout << "Mask Period: " << Lx << " x " << Ly << "\nMask Size: " << maskSizeX << " x "
       << maskSizeY << "\nMask Type: " << maskType << "\n";
  if (maskType == "LineSpace")
    cout << "Line Width: " << lineWidth << "\nSpace Width: " << spaceWidth << "\n";
  cout << "Nx: " << Nx << ",  Ny: " << Ny << endl;
  cout << "Number of Kernels: " << nk << endl << endl;

  cout << endl << "------------------------------------" << endl;
  cout << "Calculating FFT Mask..." << endl;
  cout << endl << "------------------------------------" << endl;
  cout << "Calc Image..." << endl;
  if (4 * Nx + 1 > Lx || 4 * Ny + 1 > Ly) {  // klitho_socs.cpp:164-169
    cout << "ERROR in FI()" << endl;
    exit(1);
  }
  const char* dev = getenv("KLT_DEVICE");
  klt_ctx* ctx = nullptr;
  check(klt_ctx_create(dev && *dev ? stoi(dev) : 0, &ctx));
  klt_kernels* kb = nullptr;
  if (nk > 0)
    check(klt_kernels_create(ctx, kr.data(), ki.data(), scales.data(), nk, krnSizeX, krnSizeY,
                             KLT_HOST_PTRS, &kb));
  vector<float> img((size_t)Lx * Ly, 0.0f);
  if (kb) {
    const char* path = getenv("KLT_SOCS_PATH");
    if (path && string(path) == "plain")
      check(klt_socs_image(ctx, kb, msk.data(), Lx, Ly, img.data(), KLT_HOST_PTRS));
    else
      check(klt_socs_forward(ctx, kb, msk.data(), 1, Lx, Ly, img.data(), KLT_HOST_PTRS));
    klt_kernels_destroy(kb);
  }
  klt_ctx_destroy(ctx);
  cout << "Max. Intensity: " << max_of(img) << endl;
  check(klt_write_floats("./out/image.bin", img.data(), Lx * Ly));
  check(klt_write_png_continuous("./out/image.png", Lx, Ly, img.data(), min_of(img),
                                 max_of(img)));
  check(klt_write_png_7colors("./out/image_s.png", Lx, Ly, img.data(), 0.0f,
                              KLT_TARGET_INTENSITY));

  auto dur = chrono::system_clock::now() - start;
  auto us = chrono::duration_cast<chrono::microseconds>(dur).count();
  cout << endl << "------------------------------------" << endl;
  cout << "Simulation completed." << endl;
  cout << "Total Execution Time: " << us * 1.0e-06 << " sec." << endl;
  cout << "************************************" << endl;
  return 0;
}
