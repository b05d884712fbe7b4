// klt_pupil.h -- the pupil function of the reference, shared by the TCC build
// (tcc.hip) and the Abbe source gradient (socs_fused.hip).  pupil_value turns
// floating-point contraction off for its own body, so its float arithmetic
// rounds like the reference's whatever the including file's flags.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace klt {

struct PupilGeom {
  int sh, Nx, Ny, T;
  float dz, k, Lxn, Lyn, NA, nm2;
};

struct SrcPoint {
  int p, q;
  float S;
};

// Pupil value at frequency (nx, ny) for source point (p, q), as the loops at
// klitho_tcc.cpp:202-222 compute it (returns false when the reference leaves
// the entry zero).  Zero defocus gives a unit pupil without evaluating the
// root, which extends the formula to NA > 1 (where the reference's
// sqrt(1 - rho^2 NA^2) is NaN); nm2 is the squared immersion index (1 in the
// reference).
__device__ inline bool pupil_value(const PupilGeom& g, int p, int q, int nx, int ny, float2* out) {
  // (plain operators: the pragma governs only the expressions written here)
#pragma clang fp contract(off)
  const float sx = (float)p / (float)g.sh;
  const float sy = (float)q / (float)g.sh;
  const int nxMin = (int)((-1.0f - sx) * g.Lxn);
  const int nxMax = (int)((1.0f - sx) * g.Lxn);
  const int nyMin = (int)((-1.0f - sy) * g.Lyn);
  const int nyMax = (int)((1.0f - sy) * g.Lyn);
  if (nx < nxMin || nx > nxMax || ny < nyMin || ny > nyMax) return false;
  const float fy = (float)ny / g.Lyn + sy;
  if (!(fy * fy <= 1.0f)) return false;
  const float fx = (float)nx / g.Lxn + sx;
  const float rho2 = fx * fx + fy * fy;
  if (!(rho2 <= 1.0f)) return false;
  if (g.dz == 0.0f) {
    *out = make_float2(1.0f, 0.0f);
    return true;
  }
  const float rn = rho2 * g.NA * g.NA;
  const float dzk = g.dz * g.k;
  const double root = __dsqrt_rn((double)g.nm2 - (double)rn);
  const float phi = (float)((double)dzk * root);
  *out = make_float2(cosf(phi), sinf(phi));
  return true;
}


// Geometry constants in the reference's float arithmetic (klitho_tcc.cpp:189-197).
inline PupilGeom make_geom(int srcSize, float lambda, float NA, float defocus, float nmed,
                           int Lx, int Ly, int Nx, int Ny) {
  PupilGeom g;
  g.sh = (int)((srcSize - 1) / 2.0);
  g.Nx = Nx;
  g.Ny = Ny;
  g.T = (2 * Nx + 1) * (2 * Ny + 1);
  const float na2 = NA * NA;
  g.dz = defocus / (na2 / lambda);
  g.k = (float)(2 * M_PI / lambda);
  g.Lxn = (float)Lx * NA / lambda;
  g.Lyn = (float)Ly * NA / lambda;
  g.NA = NA;
  g.nm2 = nmed * nmed;
  return g;
}

}  // namespace klt
