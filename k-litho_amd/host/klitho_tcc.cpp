// klitho_tcc -- drop-in for K-Litho's TCC tool (K-Litho-TCC/klitho_tcc.cpp:379-589)
// running the TCC build, the Hopkins image and the eigensolver on the GPU
// through libklt.
//
// Same 22 positional arguments, messages and exit codes; writes the
// reference's output tree:
//   ./out/image.bin                    (Hopkins image, float32 Ly x Lx)
//   ./out/kernels/krn_<i>_{r,i}.bin    (float32 (2Ny+1) x (2Nx+1), index-reversed)
//   ./out/kernels/kernel_info.txt
//   ./out/{source,mask,tcc_r,tcc_i,image,image_s}.png, ./out/kernels/png/krn_<i>_{r,i}.png
//                                      (the reference's PNG views, file_io.cpp:7-135)
// Options that the reference's exact argument count cannot see come from the
// environment:
//   KLT_IMMERSION_INDEX  image-side medium index n (default 1): NA may be up to n;
//                        n = 1 is the reference's defocus formula bit for bit.
//   KLT_TCC_PRECISION    "fp64" (default) or "split16" (FP16 hi+lo MFMA Gram).
//   KLT_DEVICE           GPU index (default 0).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "klt.h"

using namespace std;

namespace {

// getMax / getMin (klitho_tcc.cpp:16-35)
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

float env_float(const char* name, float dflt) {
  const char* s = getenv(name);
  return s && *s ? stof(s) : dflt;
}

[[noreturn]] void die(const string& msg) {
  cerr << msg << endl;
  exit(1);
}

void check(int rc) {
  if (rc != KLT_OK) die(klt_last_error());
}

}  // namespace

int main(int argc, char* argv[]) {
  cout << "************************************" << endl;
  cout << "* Lithography Simulation initiated..." << endl;
  cout << "************************************" << endl;
  auto start = chrono::system_clock::now();
  if (argc != 23) {
    cerr << "Error: Invalid number of arguments: " << argc << ". Expected 23 arguments." << endl;
    return 1;
  }
  const int Lx = stoi(argv[1]), Ly = stoi(argv[2]);
  const int maskSizeX = stoi(argv[3]), maskSizeY = stoi(argv[4]);
  const string maskType = argv[5];
  const int lineWidth = stoi(argv[6]), spaceWidth = stoi(argv[7]);
  const int isHorizontal = stoi(argv[8]);
  const string maskInFile = argv[9];
  const int srcSize = stoi(argv[10]);
  const string srcType = argv[11];
  const float innerR = stof(argv[12]), outerR = stof(argv[13]);
  const float radius = stof(argv[14]), offset = stof(argv[15]);
  const bool onXAxis = static_cast<bool>(stoi(argv[16]));
  const float ptX = stof(argv[17]), ptY = stof(argv[18]);
  const string sourceInFile = argv[19];
  const double NA = stof(argv[20]);
  const float defocus = stof(argv[21]);
  int nk = stoi(argv[22]);

  if (Lx <= 0 || Ly <= 0 || maskSizeX <= 0 || maskSizeY <= 0 || Lx < maskSizeX ||
      Ly < maskSizeY || srcSize <= 0) {
    cerr << "Error: Invalid input parameters. Please check the values." << endl;
    return 1;
  }
  const float wavelength = 193, lambda = wavelength, dose = 1.0f;
  const float nmed = env_float("KLT_IMMERSION_INDEX", 1.0f);
  if (NA < 0 || NA > nmed) {
    if (nmed == 1.0f)
      cerr << "Error: NA must be in the range [0, 1]." << endl;
    else
      cerr << "Error: NA must be in the range [0, " << nmed << "]." << endl;
    return 1;
  }
  const char* prec_env = getenv("KLT_TCC_PRECISION");
  const int precision =
      (prec_env && string(prec_env) == "split16") ? KLT_TCC_SPLIT_F16 : KLT_TCC_FP64;
  const int device = (int)env_float("KLT_DEVICE", 0.0f);

  vector<float> src((size_t)srcSize * srcSize, 0.0f);
  stringstream ss_src;
  ss_src << "Source Size: " << srcSize << " x " << srcSize << "\nSource Type: " << srcType << "\n";
  if (srcType == "Annular")
    ss_src << "Inner Radius: " << innerR << "\nOuter Radius: " << outerR << "\n";
  else if (srcType == "Dipole")
    ss_src << "Radius: " << radius << "\nOffset: " << offset << "\nOnXAxis: " << onXAxis << "\n";
  else if (srcType == "CrossQuadrupole")
    ss_src << "Radius: " << radius << "\nOffset: " << offset << "\n";
  else if (srcType == "Point")
    ss_src << "Point Position: (" << ptX << ", " << ptY << ")\n";
  else if (srcType == "Import")
    cout << "Read Source: " << sourceInFile << endl;
  check(klt_source_create(src.data(), srcSize, srcType.c_str(), innerR, outerR, radius, offset,
                          onXAxis ? 1 : 0, ptX, ptY, sourceInFile.c_str()));
  check(klt_write_png_gray("./out/source.png", srcSize, srcSize, src.data(), min_of(src),
                           max_of(src)));

  vector<float> msk((size_t)Lx * Ly, 0.0f);
  stringstream ss_msk;
  ss_msk << "Mask Period: " << Lx << " x " << Ly << "\nMask Size: " << maskSizeX << " x "
         << maskSizeY << "\nMask Type: " << maskType << "\n";
  if (maskType == "LineSpace")
    ss_msk << "Line Width: " << lineWidth << "\nSpace Width: " << spaceWidth << "\n";
  else if (maskType == "Import")
    cout << "Read Mask: " << maskInFile << endl;
  check(klt_mask_create(msk.data(), Lx, Ly, maskSizeX, maskSizeY, maskType.c_str(), lineWidth,
                        spaceWidth, isHorizontal, maskInFile.c_str(), dose));
  check(klt_write_png_gray("./out/mask.png", Lx, Ly, msk.data(), min_of(msk), max_of(msk)));

  float outerSigma = klt_outer_sigma(src.data(), srcSize);
  if (outerSigma > 1) {
    cout << "Warning: Outer Sigma is greater than 1. Setting outerSigma = 1." << endl;
    outerSigma = 1;
  }
  int Nx = 0, Ny = 0;
  check(klt_tcc_sizing((float)NA, Lx, Ly, outerSigma, lambda, &Nx, &Ny));
  const int tccSize = (2 * Nx + 1) * (2 * Ny + 1);
  if (tccSize < nk) {
    cout << "Warning: Number of kernels (nk) is greater than TCC size (tccSize). Setting nk = "
            "tccSize."
         << endl;
    nk = tccSize;
  }
  cout << endl << "Simulation Parameters:" << endl;
  cout << "-------------------" << endl;
  cout << ss_msk.str();
  cout << ss_src.str();
  cout << "Wavelength: " << wavelength << " nm" << endl;
  cout << "NA: " << NA << endl;
  cout << "Defocus: " << defocus << endl;
  cout << "Dose: " << dose << endl;
  cout << "Nx: " << Nx << ", Ny: " << Ny << endl;
  cout << "TCC Size: " << tccSize << " x " << tccSize << endl;
  cout << "Number of Kernels: " << nk << endl << endl;

  klt_ctx* ctx = nullptr;
  check(klt_ctx_create(device, &ctx));

  cout << endl << "------------------------------------" << endl;
  cout << "Calculating FFT Mask..." << endl;
  cout << endl << "------------------------------------" << endl;
  cout << "Starting calcTCC..." << endl;
  vector<double> tcc((size_t)tccSize * tccSize * 2);
  check(klt_tcc_build(ctx, src.data(), srcSize, outerSigma, lambda, (float)NA, defocus, nmed, Lx,
                      Ly, Nx, Ny, precision, tcc.data(), KLT_HOST_PTRS));
  {
    // complexToFloat, getMax / getMin and the two PNG views (klitho_tcc.cpp:518-529)
    vector<float> tr((size_t)tccSize * tccSize), ti((size_t)tccSize * tccSize);
    for (size_t i = 0; i < tr.size(); ++i) {
      tr[i] = (float)tcc[2 * i];
      ti[i] = (float)tcc[2 * i + 1];
    }
    cout << "Max. TCC Real Value: " << max_of(tr) << endl;
    check(klt_write_png_continuous("./out/tcc_r.png", tccSize, tccSize, tr.data(), min_of(tr),
                                   max_of(tr)));
    cout << "Max. TCC Imag Value: " << max_of(ti) << endl;
    check(klt_write_png_continuous("./out/tcc_i.png", tccSize, tccSize, ti.data(), min_of(ti),
                                   max_of(ti)));
  }

  cout << endl << "------------------------------------" << endl;
  cout << "Starting calcImage..." << endl;
  vector<float> img((size_t)Lx * Ly);
  check(klt_hopkins_image(ctx, msk.data(), Lx, Ly, tcc.data(), Nx, Ny, img.data(),
                          KLT_HOST_PTRS));
  cout << "Max. Intensity: " << max_of(img) << endl;
  check(klt_write_floats("./out/image.bin", img.data(), Lx * Ly));
  check(klt_write_png_continuous("./out/image.png", Lx, Ly, img.data(), min_of(img),
                                 max_of(img)));
  check(klt_write_png_7colors("./out/image_s.png", Lx, Ly, img.data(), 0.0f,
                              KLT_TARGET_INTENSITY));

  if (nk > 0) {
    cout << endl << "------------------------------------" << endl;
    cout << "Starting calcKernels..." << endl;
    vector<float> kr((size_t)nk * tccSize), ki((size_t)nk * tccSize), scales(nk);
    cout << "Solving eigenproblem..." << endl;
    // a non-negative source gives a positive semidefinite TCC (P^H diag(S) P)
    const unsigned psd = min_of(src) >= 0.0f ? KLT_EIG_PSD : 0u;
    const int rc = klt_calc_kernels(ctx, tcc.data(), tccSize, nk, kr.data(), ki.data(),
                                    scales.data(), KLT_HOST_PTRS | psd);
    if (rc != KLT_OK) {
      cerr << "Kernels derivation failed with info = " << rc << " (" << klt_last_error() << ")"
           << endl;
      return 1;
    }
    cout << "Eigenvalues: " << endl;
    for (int i = 0; i < nk; ++i) cout << scales[i] << " ";
    cout << endl;
    for (int i = 0; i < nk; ++i) {
      check(klt_write_floats(("./out/kernels/krn_" + to_string(i) + "_r.bin").c_str(),
                             kr.data() + (size_t)i * tccSize, tccSize));
      check(klt_write_floats(("./out/kernels/krn_" + to_string(i) + "_i.bin").c_str(),
                             ki.data() + (size_t)i * tccSize, tccSize));
      const vector<float> r(kr.begin() + (size_t)i * tccSize, kr.begin() + (size_t)(i + 1) * tccSize);
      const vector<float> im(ki.begin() + (size_t)i * tccSize, ki.begin() + (size_t)(i + 1) * tccSize);
      check(klt_write_png_continuous(("./out/kernels/png/krn_" + to_string(i) + "_r.png").c_str(),
                                     2 * Nx + 1, 2 * Ny + 1, r.data(), min_of(r), max_of(r)));
      check(klt_write_png_continuous(("./out/kernels/png/krn_" + to_string(i) + "_i.png").c_str(),
                                     2 * Nx + 1, 2 * Ny + 1, im.data(), min_of(im), max_of(im)));
    }
    check(klt_write_kernel_info("./out/kernels/kernel_info.txt", 2 * Nx + 1, 2 * Ny + 1, nk,
                                scales.data()));
  }
  klt_ctx_destroy(ctx);

  auto dur = chrono::system_clock::now() - start;
  auto us = chrono::duration_cast<chrono::microseconds>(dur).count();
  cout << endl << "------------------------------------" << endl;
  cout << "Simulation completed." << endl;
  cout << "Total Execution Time: " << us * 1.0e-06 << " sec." << endl;
  cout << "************************************" << endl;
  return 0;
}
