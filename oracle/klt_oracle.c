/*
 * klt_oracle.c -- FP64 CPU restatement of K-Litho's imaging path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU library is
 * compared against: only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline may load it.  The product (libklt.so, the klitho_* tools) never
 * links or calls it.
 *
 * Every function cites the reference lines it restates.  Discrete decisions
 * the reference takes in float (source radii, pupil membership, sizing) are
 * repeated in float with the same operation order; this file is compiled with
 * -ffp-contract=off so that no multiply-add is fused.  Everything that the
 * reference accumulates in double is accumulated in double here.
 *
 * Conventions (row-major everywhere):
 *   complex arrays are interleaved double (re, im);
 *   frequency index ID(nx,ny) = (ny+Ny)*(2Nx+1) + (nx+Nx)
 *       (K-Litho-TCC/klitho_tcc.cpp:221,245);
 *   mask spectrum centre c_in  = floor(L/2)   (klitho_socs.cpp:71-121);
 *   image synthesis centre c_out = (L+1)/2    (klitho_socs.cpp:155,192).
 * For even L both centres equal L/2.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define TWO_PI 6.283185307179586476925286766559

/* ------------------------------------------------------------------------ */
/* Sources: K-Litho-TCC/source.cpp                                           */
/* ------------------------------------------------------------------------ */

/* source.cpp:5-24 (createAnnular) */
void or_src_annular(float* m, int n, float inner, float outer) {
  int half = (n - 1) / 2;
  int ri = (int)(inner * (float)half);
  int ro = (int)(outer * (float)half);
  int c = half;
  for (int y = 0; y < n; ++y)
    for (int x = 0; x < n; ++x) {
      double d = sqrt(pow((double)(x - c), 2.0) + pow((double)(y - c), 2.0));
      m[y * n + x] = (d >= ri && d <= ro) ? 1.0f : 0.0f;
    }
}

/* source.cpp:25-57 (createCrossQuadrupole) */
void or_src_crossquad(float* m, int n, float radius, float offset) {
  int half = (n - 1) / 2;
  int off = (int)(offset * (float)half);
  int rad = (int)(radius * (float)half);
  int cx = half, cy = half;
  int cen[4][2] = {{cx, cy + off}, {cx, cy - off}, {cx + off, cy}, {cx - off, cy}};
  for (int y = 0; y < n; ++y)
    for (int x = 0; x < n; ++x) {
      int in = 0;
      for (int i = 0; i < 4 && !in; ++i) {
        double dx = x - cen[i][0], dy = y - cen[i][1];
        if (sqrt(dx * dx + dy * dy) <= rad) in = 1;
      }
      m[y * n + x] = in ? 1.0f : 0.0f;
    }
}

/* source.cpp:58-88 (createDipole); note the centre is n/2, not (n-1)/2 */
void or_src_dipole(float* m, int n, float radius, float offset, int onx) {
  int half = (n - 1) / 2;
  int off = (int)(offset * (float)half);
  int rad = (int)(radius * (float)half);
  int cx = n / 2, cy = n / 2;
  int cen[2][2] = {{cx + (onx ? off : 0), cy + (onx ? 0 : off)},
                   {cx - (onx ? off : 0), cy - (onx ? 0 : off)}};
  for (int y = 0; y < n; ++y)
    for (int x = 0; x < n; ++x) {
      int in = 0;
      for (int i = 0; i < 2 && !in; ++i) {
        double dx = x - cen[i][0], dy = y - cen[i][1];
        if (sqrt(dx * dx + dy * dy) <= rad) in = 1;
      }
      m[y * n + x] = in ? 1.0f : 0.0f;
    }
}

/* source.cpp:122-129 (Point branch of createSource); m must be zeroed */
void or_src_point(float* m, int n, float px, float py) {
  float fx = px * (float)n;
  float fy = (-1.0f * py) * (float)n;
  int ix = (int)round(fx / 2.0 + n / 2.0);
  int iy = (int)round(fy / 2.0 + n / 2.0);
  if (px >= 0) ix -= 1;
  if (-1.0f * py >= 0) iy -= 1;
  if (ix >= 0 && ix < n && iy >= 0 && iy < n) m[iy * n + ix] = 1.0f;
}

/* source.cpp:89-103 (normalizeMatrix): sequential float sum */
void or_normalize(float* m, int count) {
  float s = 0.0f;
  for (int i = 0; i < count; ++i) s += m[i];
  if (s != 0.0f)
    for (int i = 0; i < count; ++i) m[i] /= s;
}

/* ------------------------------------------------------------------------ */
/* Masks: common/mask.cpp                                                    */
/* ------------------------------------------------------------------------ */

/* mask.cpp:5-29 (generateLineSpace) */
void or_mask_linespace(float* m, int sx, int sy, int lw, int sw, int horiz) {
  int period = lw + sw;
  for (int y = 0; y < sy; ++y)
    for (int x = 0; x < sx; ++x) {
      int t = horiz ? (y % period) : (x % period);
      m[y * sx + x] = (t < lw) ? 1.0f : 0.0f;
    }
}

/* mask.cpp:53-59: centre a msx x msy mask in the Lx x Ly period, times dose.
 * out must be zeroed by the caller. */
void or_mask_place(const float* in, int msx, int msy, float* out, int Lx, int Ly,
                   float dose) {
  int dy = (Ly - msy) / 2, dx = (Lx - msx) / 2;
  for (int y = 0; y < msy; ++y)
    for (int x = 0; x < msx; ++x) out[(y + dy) * Lx + x + dx] = in[y * msx + x] * dose;
}

/* ------------------------------------------------------------------------ */
/* Sizing: klitho_tcc.cpp:52-67, 479-491                                     */
/* ------------------------------------------------------------------------ */

/* klitho_tcc.cpp:52-67 (getOuterSigma), unclamped */
float or_outer_sigma(const float* src, int n) {
  int c = (n - 1) / 2;
  float maxd = 0.0f;
  for (int y = 0; y < n; ++y)
    for (int x = 0; x < n; ++x)
      if (src[y * n + x] != 0.0f) {
        float d = (float)sqrt((double)((y - c) * (y - c) + (x - c) * (x - c)));
        if (d > maxd) maxd = d;
      }
  return maxd / (float)c;
}

/* klitho_tcc.cpp:479-486: clamp sigma, Nx = floor(f32(NA)*Lx*(1+sigma)/lambda) */
void or_sizing(float NA, int Lx, int Ly, float sigma, float lambda, int* Nx, int* Ny) {
  if (sigma > 1.0f) sigma = 1.0f;
  float one_s = 1.0f + sigma;
  float ax = NA * (float)Lx;
  ax = ax * one_s;
  ax = ax / lambda;
  float ay = NA * (float)Ly;
  ay = ay * one_s;
  ay = ay / lambda;
  *Nx = (int)floorf(ax);
  *Ny = (int)floorf(ay);
}

/* ------------------------------------------------------------------------ */
/* Pupil and TCC: klitho_tcc.cpp:187-267                                     */
/* ------------------------------------------------------------------------ */

typedef struct {
  int sh, oSgm;
  float dz, k, Lxn, Lyn, NA, nm2;
} pupil_geom;

/* nmed: refractive index of the image-side medium.  nmed = 1 is the
 * reference's formula bit for bit; nmed > 1 extends it to immersion
 * (NA > 1), where the reference's sqrt(1 - rho^2 NA^2) is NaN. */
static void geom_init(pupil_geom* g, int srcSize, float outerSigma, float lambda,
                      float NA, float defocus, int Lx, int Ly, float nmed) {
  g->sh = (int)((srcSize - 1) / 2.0);                  /* :189 */
  g->oSgm = (int)ceilf((float)g->sh * outerSigma);     /* :190 */
  float na2 = NA * NA;
  g->dz = defocus / (na2 / lambda);                    /* :191 */
  g->k = (float)(2 * 3.14159265358979323846 / lambda); /* :191, M_PI in double */
  g->Lxn = (float)Lx * NA / lambda;                    /* :197 */
  g->Lyn = (float)Ly * NA / lambda;
  g->NA = NA;
  g->nm2 = nmed * nmed;
}

/* Pupil value for source point (p,q) at frequency (nx,ny).  Returns 1 and
 * writes (re,im) when the reference's loops at klitho_tcc.cpp:206-222 visit
 * the frequency and it passes both membership tests, else 0. */
static int pupil_at(const pupil_geom* g, int p, int q, int nx, int ny, double* re,
                    double* im) {
  float sx = (float)p / (float)g->sh;                  /* :206 */
  float sy = (float)q / (float)g->sh;
  int nxMin = (int)((-1.0f - sx) * g->Lxn);            /* :208-211 */
  int nxMax = (int)((1.0f - sx) * g->Lxn);
  int nyMin = (int)((-1.0f - sy) * g->Lyn);
  int nyMax = (int)((1.0f - sy) * g->Lyn);
  if (nx < nxMin || nx > nxMax || ny < nyMin || ny > nyMax) return 0;
  float fy = (float)ny / g->Lyn + sy;                  /* :214 */
  if (!(fy * fy <= 1.0f)) return 0;                    /* :215 */
  float fx = (float)nx / g->Lxn + sx;                  /* :217 */
  float rho2 = fx * fx + fy * fy;                      /* :218 */
  if (!(rho2 <= 1.0)) return 0;                        /* :219 */
  float rn = rho2 * g->NA;
  rn = rn * g->NA;
  float dzk = g->dz * g->k;
  if (g->dz == 0.0f) { /* zero defocus: unit pupil, no root evaluated */
    *re = 1.0;
    *im = 0.0;
    return 1;
  }
  float phi = (float)((double)dzk * sqrt((double)g->nm2 - (double)rn)); /* :220 */
  *re = (double)cosf(phi);                             /* :221 */
  *im = (double)sinf(phi);
  return 1;
}

/* Number of source points the reference uses (S != 0, inside the unit
 * disc, |p|,|q| <= oSgm), klitho_tcc.cpp:202-205. */
int or_count_source_points(const float* src, int srcSize, float outerSigma) {
  int sh = (int)((srcSize - 1) / 2.0);
  int oS = (int)ceilf((float)sh * outerSigma);
  int cnt = 0;
  for (int q = -oS; q <= oS; ++q)
    for (int p = -oS; p <= oS; ++p) {
      int id = (q + sh) * srcSize + p + sh;
      if (src[id] != 0.0f && p * p + q * q <= sh * sh) ++cnt;
    }
  return cnt;
}

/* Pupil matrix over the used source points, in the reference's loop order
 * (q outer, p inner).  pup: [nsrc][T] complex (zeros where outside);
 * sval: [nsrc] source weights; pq: [nsrc][2] (p,q).  Returns nsrc. */
int or_pupil_matrix(const float* src, int srcSize, float outerSigma, float lambda,
                    float NA, float defocus, int Lx, int Ly, int Nx, int Ny, float nmed,
                    double* pup, double* sval, int* pq) {
  pupil_geom g;
  geom_init(&g, srcSize, outerSigma, lambda, NA, defocus, Lx, Ly, nmed);
  int W = 2 * Nx + 1, T = W * (2 * Ny + 1);
  int s = 0;
  for (int q = -g.oSgm; q <= g.oSgm; ++q)
    for (int p = -g.oSgm; p <= g.oSgm; ++p) {
      int id = (q + g.sh) * srcSize + p + g.sh;
      if (!(src[id] != 0.0f && p * p + q * q <= g.sh * g.sh)) continue;
      double* row = pup + (size_t)s * T * 2;
      memset(row, 0, sizeof(double) * 2 * T);
      for (int ny = -Ny; ny <= Ny; ++ny)
        for (int nx = -Nx; nx <= Nx; ++nx) {
          double re, im;
          if (pupil_at(&g, p, q, nx, ny, &re, &im)) {
            int ID = (ny + Ny) * W + nx + Nx;
            row[2 * ID] = re;
            row[2 * ID + 1] = im;
          }
        }
      sval[s] = (double)src[id];
      pq[2 * s] = p;
      pq[2 * s + 1] = q;
      ++s;
    }
  return s;
}

/* klitho_tcc.cpp:187-267 (calcTCC): TCC[i][j] = sum_s S_s P(i,s) conj(P(j,s)),
 * accumulated in double over s in the reference's order; upper triangle
 * computed, lower triangle its conjugate.  tcc: T*T complex, zeroed here. */
void or_tcc(double* tcc, const float* src, int srcSize, float outerSigma,
            float lambda, float NA, float defocus, int Lx, int Ly, int Nx, int Ny,
            float nmed) {
  int W = 2 * Nx + 1, T = W * (2 * Ny + 1);
  int ns = or_count_source_points(src, srcSize, outerSigma);
  double* pup = (double*)malloc(sizeof(double) * 2 * (size_t)T * (ns > 0 ? ns : 1));
  double* sv = (double*)malloc(sizeof(double) * (ns > 0 ? ns : 1));
  int* pq = (int*)malloc(sizeof(int) * 2 * (ns > 0 ? ns : 1));
  or_pupil_matrix(src, srcSize, outerSigma, lambda, NA, defocus, Lx, Ly, Nx, Ny, nmed,
                  pup, sv, pq);
  memset(tcc, 0, sizeof(double) * 2 * (size_t)T * T);
  for (int s = 0; s < ns; ++s) {
    const double* P = pup + (size_t)s * T * 2;
    double S = sv[s];
    for (int i = 0; i < T; ++i) {
      double ar = P[2 * i], ai = P[2 * i + 1];
      if (ar == 0.0 && ai == 0.0) continue;          /* :246 */
      double tr = ar * S, ti = ai * S;               /* :247 */
      double* row = tcc + (size_t)i * T * 2;
      for (int j = i; j < T; ++j) {                  /* :248-253 */
        double br = P[2 * j], bi = -P[2 * j + 1];
        row[2 * j] += tr * br - ti * bi;
        row[2 * j + 1] += tr * bi + ti * br;
      }
    }
  }
  for (int i = 0; i < T; ++i)                        /* :262-266 */
    for (int j = i + 1; j < T; ++j) {
      tcc[2 * ((size_t)j * T + i)] = tcc[2 * ((size_t)i * T + j)];
      tcc[2 * ((size_t)j * T + i) + 1] = -tcc[2 * ((size_t)i * T + j) + 1];
    }
  free(pup);
  free(sv);
  free(pq);
}

/* ------------------------------------------------------------------------ */
/* Hermitian eigensolver (replaces zheevr_, klitho_tcc.cpp:301-377)          */
/* Householder reduction to tridiagonal form, phase scaling to a real        */
/* symmetric tridiagonal, implicit-shift QL with accumulated rotations,      */
/* back-transformation.  Textbook algorithms (Golub & Van Loan ch. 8).       */
/* ------------------------------------------------------------------------ */

static void tql_implicit(int n, double* d, double* e, double* z /* n*n, z[r*n+c] */) {
  /* e[i] couples d[i] and d[i+1], i < n-1; e[n-1] = 0 on entry. */
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    for (;;) {
      int m;
      for (m = l; m < n - 1; ++m) {
        double dd = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) <= 2.220446049250313e-16 * dd) break;
      }
      if (m == l) break;
      if (++iter > 200) break;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0 ? fabs(r) : -fabs(r)));
      double s = 1.0, c = 1.0, p = 0.0;
      int i;
      int early = 0;
      for (i = m - 1; i >= l; --i) {
        double f = s * e[i], b = c * e[i];
        r = hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[m] = 0.0;
          early = 1;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        for (int k = 0; k < n; ++k) {
          double zf = z[(size_t)k * n + i + 1];
          double zi = z[(size_t)k * n + i];
          z[(size_t)k * n + i + 1] = s * zi + c * zf;
          z[(size_t)k * n + i] = c * zi - s * zf;
        }
      }
      if (early) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
}

/* Top-k eigenpairs of the Hermitian matrix a (n x n complex, row-major; only
 * the upper triangle is read, as zheevr_ with uplo='U' at klitho_tcc.cpp:340
 * reads it).  w: k eigenvalues, descending.  z: k x n complex, row i is the
 * unit eigenvector of w[i].  Returns 0 on success. */
int or_eigh_topk(const double* a_in, int n, int k, double* w, double* z) {
  if (k > n) k = n;
  size_t nn = (size_t)n * n;
  double* a = (double*)malloc(sizeof(double) * 2 * nn);
  /* Hermitian full storage from the upper triangle */
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      if (j >= i) {
        a[2 * ((size_t)i * n + j)] = a_in[2 * ((size_t)i * n + j)];
        a[2 * ((size_t)i * n + j) + 1] = a_in[2 * ((size_t)i * n + j) + 1];
      } else {
        a[2 * ((size_t)i * n + j)] = a_in[2 * ((size_t)j * n + i)];
        a[2 * ((size_t)i * n + j) + 1] = -a_in[2 * ((size_t)j * n + i) + 1];
      }
    }
  for (int i = 0; i < n; ++i) a[2 * ((size_t)i * n + i) + 1] = 0.0;
  /* Householder vectors v_j (length n, zero above j+1) and taus */
  double* V = (double*)calloc(2 * nn, sizeof(double));
  double* tau = (double*)calloc(n, sizeof(double));
  double* pv = (double*)malloc(sizeof(double) * 2 * n);
  double* wv = (double*)malloc(sizeof(double) * 2 * n);
  for (int j = 0; j + 2 < n; ++j) {
    /* x = A[j+1:, j] */
    double xnorm2 = 0.0;
    for (int i = j + 1; i < n; ++i) {
      double xr = a[2 * ((size_t)i * n + j)], xi = a[2 * ((size_t)i * n + j) + 1];
      xnorm2 += xr * xr + xi * xi;
    }
    double x0r = a[2 * ((size_t)(j + 1) * n + j)], x0i = a[2 * ((size_t)(j + 1) * n + j) + 1];
    double tail = xnorm2 - (x0r * x0r + x0i * x0i);
    if (tail <= 0.0) { tau[j] = 0.0; continue; }
    double xnorm = sqrt(xnorm2);
    double ax0 = hypot(x0r, x0i);
    double phr = 1.0, phi = 0.0;
    if (ax0 > 0.0) { phr = x0r / ax0; phi = x0i / ax0; }
    /* v = x + e^{i theta} |x| e_1 ; H = I - tau v v^H, tau = 2 / (v^H v) */
    double* v = V + 2 * (size_t)j * n;
    for (int i = j + 1; i < n; ++i) {
      v[2 * i] = a[2 * ((size_t)i * n + j)];
      v[2 * i + 1] = a[2 * ((size_t)i * n + j) + 1];
    }
    v[2 * (j + 1)] += phr * xnorm;
    v[2 * (j + 1) + 1] += phi * xnorm;
    double vn2 = 0.0;
    for (int i = j + 1; i < n; ++i) vn2 += v[2 * i] * v[2 * i] + v[2 * i + 1] * v[2 * i + 1];
    double t = 2.0 / vn2;
    tau[j] = t;
    /* p = t * A22 v ; w = p - (t/2)(v^H p) v ; A22 -= v w^H + w v^H */
    for (int i = j + 1; i < n; ++i) {
      double sr = 0.0, si = 0.0;
      const double* ar = a + 2 * (size_t)i * n;
      for (int l = j + 1; l < n; ++l) {
        double mr = ar[2 * l], mi = ar[2 * l + 1];
        sr += mr * v[2 * l] - mi * v[2 * l + 1];
        si += mr * v[2 * l + 1] + mi * v[2 * l];
      }
      pv[2 * i] = t * sr;
      pv[2 * i + 1] = t * si;
    }
    double hr = 0.0, hi = 0.0; /* v^H p */
    for (int i = j + 1; i < n; ++i) {
      hr += v[2 * i] * pv[2 * i] + v[2 * i + 1] * pv[2 * i + 1];
      hi += v[2 * i] * pv[2 * i + 1] - v[2 * i + 1] * pv[2 * i];
    }
    double cr = 0.5 * t * hr, ci = 0.5 * t * hi;
    for (int i = j + 1; i < n; ++i) {
      wv[2 * i] = pv[2 * i] - (cr * v[2 * i] - ci * v[2 * i + 1]);
      wv[2 * i + 1] = pv[2 * i + 1] - (cr * v[2 * i + 1] + ci * v[2 * i]);
    }
    for (int i = j + 1; i < n; ++i) {
      double vir = v[2 * i], vii = v[2 * i + 1], wir = wv[2 * i], wii = wv[2 * i + 1];
      double* ar = a + 2 * (size_t)i * n;
      for (int l = j + 1; l < n; ++l) {
        /* v_i conj(w_l) + w_i conj(v_l) */
        double wlr = wv[2 * l], wli = wv[2 * l + 1], vlr = v[2 * l], vli = v[2 * l + 1];
        ar[2 * l] -= (vir * wlr + vii * wli) + (wir * vlr + wii * vli);
        ar[2 * l + 1] -= (vii * wlr - vir * wli) + (wii * vlr - wir * vli);
      }
    }
    /* column j below the subdiagonal becomes -e^{i theta}|x| e_1 */
    a[2 * ((size_t)(j + 1) * n + j)] = -phr * xnorm;
    a[2 * ((size_t)(j + 1) * n + j) + 1] = -phi * xnorm;
    a[2 * ((size_t)j * n + j + 1)] = -phr * xnorm;
    a[2 * ((size_t)j * n + j + 1) + 1] = phi * xnorm;
    for (int i = j + 2; i < n; ++i) {
      a[2 * ((size_t)i * n + j)] = a[2 * ((size_t)i * n + j) + 1] = 0.0;
      a[2 * ((size_t)j * n + i)] = a[2 * ((size_t)j * n + i) + 1] = 0.0;
    }
  }
  /* tridiagonal: d real, off-diagonal e_c[i] = A[i+1][i] complex */
  double* d = (double*)malloc(sizeof(double) * n);
  double* e = (double*)calloc(n, sizeof(double));
  double* dr = (double*)malloc(sizeof(double) * n);
  double* di = (double*)malloc(sizeof(double) * n);
  for (int i = 0; i < n; ++i) d[i] = a[2 * ((size_t)i * n + i)];
  dr[0] = 1.0;
  di[0] = 0.0;
  for (int i = 0; i + 1 < n; ++i) {
    double er = a[2 * ((size_t)(i + 1) * n + i)], ei = a[2 * ((size_t)(i + 1) * n + i) + 1];
    double ab = hypot(er, ei);
    e[i] = ab;
    double ur = 1.0, ui = 0.0;
    if (ab > 0.0) { ur = er / ab; ui = ei / ab; }
    /* delta_{i+1} = delta_i * e_i/|e_i| */
    dr[i + 1] = dr[i] * ur - di[i] * ui;
    di[i + 1] = dr[i] * ui + di[i] * ur;
  }
  double* Z = (double*)calloc(nn, sizeof(double));
  for (int i = 0; i < n; ++i) Z[(size_t)i * n + i] = 1.0;
  tql_implicit(n, d, e, Z);
  /* select top-k by value */
  int* idx = (int*)malloc(sizeof(int) * n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  for (int i = 0; i < k; ++i) {
    int best = i;
    for (int j = i + 1; j < n; ++j)
      if (d[idx[j]] > d[idx[best]]) best = j;
    int t = idx[i];
    idx[i] = idx[best];
    idx[best] = t;
  }
  for (int c = 0; c < k; ++c) {
    int col = idx[c];
    w[c] = d[col];
    double* y = z + 2 * (size_t)c * n;
    for (int i = 0; i < n; ++i) {
      double zz = Z[(size_t)i * n + col];
      y[2 * i] = dr[i] * zz;
      y[2 * i + 1] = di[i] * zz;
    }
    /* x = H_0 H_1 ... H_{n-3} y */
    for (int j = n - 3; j >= 0; --j) {
      if (tau[j] == 0.0) continue;
      const double* v = V + 2 * (size_t)j * n;
      double sr = 0.0, si = 0.0; /* v^H y */
      for (int i = j + 1; i < n; ++i) {
        sr += v[2 * i] * y[2 * i] + v[2 * i + 1] * y[2 * i + 1];
        si += v[2 * i] * y[2 * i + 1] - v[2 * i + 1] * y[2 * i];
      }
      sr *= tau[j];
      si *= tau[j];
      for (int i = j + 1; i < n; ++i) {
        y[2 * i] -= v[2 * i] * sr - v[2 * i + 1] * si;
        y[2 * i + 1] -= v[2 * i] * si + v[2 * i + 1] * sr;
      }
    }
  }
  free(a); free(V); free(tau); free(pv); free(wv); free(d); free(e);
  free(dr); free(di); free(Z); free(idx);
  return 0;
}

/* klitho_tcc.cpp:343-362: kernels are the eigenvectors of the k largest
 * eigenvalues, index-reversed (krn[j] = z[T-1-j]) and stored as float;
 * scales are the eigenvalues as float.  vec: k x T complex (rows). */
void or_kernels_from_eig(const double* vec, const double* w, int k, int T, float* kr,
                         float* ki, float* scales) {
  for (int i = 0; i < k; ++i) {
    scales[i] = (float)w[i];
    for (int j = 0; j < T; ++j) {
      kr[(size_t)i * T + j] = (float)vec[2 * ((size_t)i * T + (T - 1 - j))];
      ki[(size_t)i * T + j] = (float)vec[2 * ((size_t)i * T + (T - 1 - j)) + 1];
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Mask spectrum band and SOCS image: klitho_socs.cpp:71-121, 196-241        */
/* ------------------------------------------------------------------------ */

/* exp(sign * 2 pi i r / L) for r in [0,L), exact integer argument reduction */
static void twiddles(int L, int sign, double* tw) {
  for (int r = 0; r < L; ++r) {
    double a = TWO_PI * (double)r / (double)L;
    tw[2 * r] = cos(a);
    tw[2 * r + 1] = sign * sin(a);
  }
}

static inline int imod(long long a, int L) {
  long long r = a % L;
  return (int)(r < 0 ? r + L : r);
}

/* M^[n] = 1/(Lx Ly) sum_x m(x) exp(-2 pi i n.(x - floor(L/2))/L), |nx|<=Bx,
 * |ny|<=By (the centred spectrum FT_r2c produces, klitho_socs.cpp:71-121).
 * mask may be float (mf) or double (md); band: (2By+1) x (2Bx+1) complex. */
static void band_dft(const float* mf, const double* md, int Lx, int Ly, int Bx, int By,
                     double* band) {
  int Wx = 2 * Bx + 1, Wy = 2 * By + 1;
  int cx = Lx / 2, cy = Ly / 2;
  double* twx = (double*)malloc(sizeof(double) * 2 * Lx);
  double* twy = (double*)malloc(sizeof(double) * 2 * Ly);
  twiddles(Lx, -1, twx);
  twiddles(Ly, -1, twy);
  double* R = (double*)calloc((size_t)Ly * Wx * 2, sizeof(double));
  for (int y = 0; y < Ly; ++y)
    for (int x = 0; x < Lx; ++x) {
      double v = mf ? (double)mf[(size_t)y * Lx + x] : md[(size_t)y * Lx + x];
      if (v == 0.0) continue;
      for (int nx = -Bx; nx <= Bx; ++nx) {
        int r = imod((long long)nx * (x - cx), Lx);
        R[2 * ((size_t)y * Wx + nx + Bx)] += v * twx[2 * r];
        R[2 * ((size_t)y * Wx + nx + Bx) + 1] += v * twx[2 * r + 1];
      }
    }
  double inv = 1.0 / ((double)Lx * (double)Ly);
  for (int ny = -By; ny <= By; ++ny)
    for (int nx = -Bx; nx <= Bx; ++nx) {
      double sr = 0.0, si = 0.0;
      for (int y = 0; y < Ly; ++y) {
        int r = imod((long long)ny * (y - cy), Ly);
        double ar = R[2 * ((size_t)y * Wx + nx + Bx)], ai = R[2 * ((size_t)y * Wx + nx + Bx) + 1];
        sr += ar * twy[2 * r] - ai * twy[2 * r + 1];
        si += ar * twy[2 * r + 1] + ai * twy[2 * r];
      }
      band[2 * ((size_t)(ny + By) * Wx + nx + Bx)] = sr * inv;
      band[2 * ((size_t)(ny + By) * Wx + nx + Bx) + 1] = si * inv;
    }
  (void)Wy;
  free(twx); free(twy); free(R);
}

void or_mask_band(const float* m, int Lx, int Ly, int Nx, int Ny, double* band) {
  band_dft(m, NULL, Lx, Ly, Nx, Ny, band);
}

/* Field E(x) = sum_n A[n] exp(+2 pi i n.(x - c_out)/L) at every pixel,
 * separable; A: (2Ny+1)x(2Nx+1) complex; E: Ly x Lx complex. */
static void synth_field(const double* A, int Nx, int Ny, int Lx, int Ly,
                        const double* twx, const double* twy, double* tmp, double* E) {
  int Wx = 2 * Nx + 1;
  int cx = (Lx + 1) / 2, cy = (Ly + 1) / 2;
  /* tmp[ny][x] = sum_nx A[ny][nx] e^{+i nx (x-cx)} */
  for (int ny = 0; ny < 2 * Ny + 1; ++ny)
    for (int x = 0; x < Lx; ++x) {
      double sr = 0.0, si = 0.0;
      for (int nx = -Nx; nx <= Nx; ++nx) {
        int r = imod((long long)nx * (x - cx), Lx);
        double ar = A[2 * (ny * Wx + nx + Nx)], ai = A[2 * (ny * Wx + nx + Nx) + 1];
        sr += ar * twx[2 * r] - ai * twx[2 * r + 1];
        si += ar * twx[2 * r + 1] + ai * twx[2 * r];
      }
      tmp[2 * ((size_t)ny * Lx + x)] = sr;
      tmp[2 * ((size_t)ny * Lx + x) + 1] = si;
    }
  for (int y = 0; y < Ly; ++y) {
    double* er = E + 2 * (size_t)y * Lx;
    memset(er, 0, sizeof(double) * 2 * Lx);
    for (int ny = -Ny; ny <= Ny; ++ny) {
      int r = imod((long long)ny * (y - cy), Ly);
      double wr = twy[2 * r], wi = twy[2 * r + 1];
      const double* tr = tmp + 2 * (size_t)(ny + Ny) * Lx;
      for (int x = 0; x < Lx; ++x) {
        er[2 * x] += tr[2 * x] * wr - tr[2 * x + 1] * wi;
        er[2 * x + 1] += tr[2 * x] * wi + tr[2 * x + 1] * wr;
      }
    }
  }
}

/* klitho_socs.cpp:196-241 (calcSOCS + FI), as the closed form the coarse
 * grid plus Fourier interpolation evaluates exactly:
 *   I(x) = sum_k scales[k] |sum_n K_k[n] M^[n] exp(+2 pi i n.(x-c_out)/L)|^2.
 * kr/ki: nk x T float kernels (file layout), img: Ly x Lx double. */
void or_socs_image(const float* mask, int Lx, int Ly, const float* kr, const float* ki,
                   const float* scales, int nk, int Nx, int Ny, double* img) {
  int Wx = 2 * Nx + 1, T = Wx * (2 * Ny + 1);
  double* band = (double*)malloc(sizeof(double) * 2 * T);
  band_dft(mask, NULL, Lx, Ly, Nx, Ny, band);
  double* twx = (double*)malloc(sizeof(double) * 2 * Lx);
  double* twy = (double*)malloc(sizeof(double) * 2 * Ly);
  twiddles(Lx, +1, twx);
  twiddles(Ly, +1, twy);
  double* A = (double*)malloc(sizeof(double) * 2 * T);
  double* tmp = (double*)malloc(sizeof(double) * 2 * (size_t)(2 * Ny + 1) * Lx);
  double* E = (double*)malloc(sizeof(double) * 2 * (size_t)Lx * Ly);
  memset(img, 0, sizeof(double) * (size_t)Lx * Ly);
  for (int k = 0; k < nk; ++k) {
    for (int i = 0; i < T; ++i) {
      double a = kr[(size_t)k * T + i], b = ki[(size_t)k * T + i];
      A[2 * i] = a * band[2 * i] - b * band[2 * i + 1];
      A[2 * i + 1] = a * band[2 * i + 1] + b * band[2 * i];
    }
    synth_field(A, Nx, Ny, Lx, Ly, twx, twy, tmp, E);
    double lam = (double)scales[k];
    for (size_t p = 0; p < (size_t)Lx * Ly; ++p)
      img[p] += lam * (E[2 * p] * E[2 * p] + E[2 * p + 1] * E[2 * p + 1]);
  }
  free(band); free(twx); free(twy); free(A); free(tmp); free(E);
}

/* klitho_tcc.cpp:270-290 (calcImage) then FT_c2r (:151-183): the Hopkins
 * image from the full TCC.  The half spectrum gx in [0,2Nx], gy in
 * [-2Ny,2Ny] is summed as a c2r transform does: the gx = 0 column through
 * its real part, the other columns twice through theirs. */
void or_hopkins_image(const float* mask, int Lx, int Ly, const double* tcc, int Nx,
                      int Ny, double* img) {
  int Wx = 2 * Nx + 1, T = Wx * (2 * Ny + 1);
  int half = (T - 1) / 2;
  double* band = (double*)malloc(sizeof(double) * 2 * T);
  band_dft(mask, NULL, Lx, Ly, Nx, Ny, band);
  int Gx = 2 * Nx + 1, Gy = 4 * Ny + 1;
  double* spec = (double*)calloc((size_t)Gx * Gy * 2, sizeof(double));
  for (int g2y = -2 * Ny; g2y <= 2 * Ny; ++g2y)
    for (int g2x = 0; g2x <= 2 * Nx; ++g2x) {
      double vr = 0.0, vi = 0.0;
      for (int n1y = -Ny; n1y <= Ny; ++n1y)
        for (int n1x = -Nx; n1x <= Nx; ++n1x) {
          if (abs(g2x + n1x) > Nx || abs(g2y + n1y) > Ny) continue;
          const double* ma = band + 2 * ((g2y + n1y + Ny) * Wx + g2x + n1x + Nx);
          const double* mb = band + 2 * ((n1y + Ny) * Wx + n1x + Nx);
          int i1 = n1y * Wx + n1x + half, i2 = (g2y + n1y) * Wx + g2x + n1x + half;
          const double* t = tcc + 2 * ((size_t)i1 * T + i2);
          /* ma * conj(mb) * t */
          double pr = ma[0] * mb[0] + ma[1] * mb[1];
          double pi = ma[1] * mb[0] - ma[0] * mb[1];
          vr += pr * t[0] - pi * t[1];
          vi += pr * t[1] + pi * t[0];
        }
      spec[2 * ((g2y + 2 * Ny) * Gx + g2x)] = vr;
      spec[2 * ((g2y + 2 * Ny) * Gx + g2x) + 1] = vi;
    }
  int cx = (Lx + 1) / 2, cy = (Ly + 1) / 2;
  double* twx = (double*)malloc(sizeof(double) * 2 * Lx);
  double* twy = (double*)malloc(sizeof(double) * 2 * Ly);
  twiddles(Lx, +1, twx);
  twiddles(Ly, +1, twy);
  /* Y[y][gx] = sum_gy spec[gy][gx] e^{+i gy (y - cy)} */
  double* Y = (double*)malloc(sizeof(double) * 2 * (size_t)Ly * Gx);
  for (int y = 0; y < Ly; ++y)
    for (int gx = 0; gx < Gx; ++gx) {
      double sr = 0.0, si = 0.0;
      for (int gy = -2 * Ny; gy <= 2 * Ny; ++gy) {
        int r = imod((long long)gy * (y - cy), Ly);
        const double* s = spec + 2 * ((gy + 2 * Ny) * Gx + gx);
        sr += s[0] * twy[2 * r] - s[1] * twy[2 * r + 1];
        si += s[0] * twy[2 * r + 1] + s[1] * twy[2 * r];
      }
      Y[2 * ((size_t)y * Gx + gx)] = sr;
      Y[2 * ((size_t)y * Gx + gx) + 1] = si;
    }
  for (int y = 0; y < Ly; ++y)
    for (int x = 0; x < Lx; ++x) {
      double acc = Y[2 * ((size_t)y * Gx)];
      for (int gx = 1; gx < Gx; ++gx) {
        int r = imod((long long)gx * (x - cx), Lx);
        const double* yy = Y + 2 * ((size_t)y * Gx + gx);
        acc += 2.0 * (yy[0] * twx[2 * r] - yy[1] * twx[2 * r + 1]);
      }
      img[(size_t)y * Lx + x] = acc;
    }
  free(band); free(spec); free(twx); free(twy); free(Y);
}

/* ------------------------------------------------------------------------ */
/* Adjoint (no reference counterpart): loss and gradients in FP64            */
/* ------------------------------------------------------------------------ */

static inline double sigm(double t) { return 1.0 / (1.0 + exp(-t)); }

/* L = sum_x (sigmoid(a (I(x) - thr)) - T(x))^2 and dL/dI. */
static double resist_loss(const double* I, const float* tgt, size_t np, double a,
                          double thr, double* G) {
  double L = 0.0;
  for (size_t p = 0; p < np; ++p) {
    double s = sigm(a * (I[p] - thr));
    double r = s - (double)tgt[p];
    L += r * r;
    if (G) G[p] = 2.0 * r * a * s * (1.0 - s);
  }
  return L;
}

/* SOCS forward with a double mask (for finite differences), img: Ly x Lx */
void or_socs_image_d(const double* mask, int Lx, int Ly, const float* kr, const float* ki,
                     const float* scales, int nk, int Nx, int Ny, double* img) {
  int Wx = 2 * Nx + 1, T = Wx * (2 * Ny + 1);
  double* band = (double*)malloc(sizeof(double) * 2 * T);
  band_dft(NULL, mask, Lx, Ly, Nx, Ny, band);
  double* twx = (double*)malloc(sizeof(double) * 2 * Lx);
  double* twy = (double*)malloc(sizeof(double) * 2 * Ly);
  twiddles(Lx, +1, twx);
  twiddles(Ly, +1, twy);
  double* A = (double*)malloc(sizeof(double) * 2 * T);
  double* tmp = (double*)malloc(sizeof(double) * 2 * (size_t)(2 * Ny + 1) * Lx);
  double* E = (double*)malloc(sizeof(double) * 2 * (size_t)Lx * Ly);
  memset(img, 0, sizeof(double) * (size_t)Lx * Ly);
  for (int k = 0; k < nk; ++k) {
    for (int i = 0; i < T; ++i) {
      double a = kr[(size_t)k * T + i], b = ki[(size_t)k * T + i];
      A[2 * i] = a * band[2 * i] - b * band[2 * i + 1];
      A[2 * i + 1] = a * band[2 * i + 1] + b * band[2 * i];
    }
    synth_field(A, Nx, Ny, Lx, Ly, twx, twy, tmp, E);
    for (size_t p = 0; p < (size_t)Lx * Ly; ++p)
      img[p] += (double)scales[k] * (E[2 * p] * E[2 * p] + E[2 * p + 1] * E[2 * p + 1]);
  }
  free(band); free(twx); free(twy); free(A); free(tmp); free(E);
}

/* Resist loss of the SOCS image and its gradient with respect to the mask:
 *   dL/dm(x') = (2/(Lx Ly)) Re sum_k l_k sum_n K_k[n] e^{-2 pi i n.(x'-c_in)/L}
 *               sum_x G(x) conj(E_k(x)) e^{+2 pi i n.(x-c_out)/L}
 * evaluated pixel by pixel (no band shortcuts).  grad may be NULL. */
double or_socs_loss_grad(const double* mask, const float* tgt, int Lx, int Ly,
                         const float* kr, const float* ki, const float* scales, int nk,
                         int Nx, int Ny, double a, double thr, double* img_out,
                         double* grad) {
  int Wx = 2 * Nx + 1, Wy = 2 * Ny + 1, T = Wx * Wy;
  size_t np = (size_t)Lx * Ly;
  double* band = (double*)malloc(sizeof(double) * 2 * T);
  band_dft(NULL, mask, Lx, Ly, Nx, Ny, band);
  double* twx = (double*)malloc(sizeof(double) * 2 * Lx);
  double* twy = (double*)malloc(sizeof(double) * 2 * Ly);
  twiddles(Lx, +1, twx);
  twiddles(Ly, +1, twy);
  double* A = (double*)malloc(sizeof(double) * 2 * T);
  double* tmp = (double*)malloc(sizeof(double) * 2 * (size_t)Wy * Lx);
  double* E = (double*)malloc(sizeof(double) * 2 * np * (nk > 0 ? nk : 1));
  double* I = img_out ? img_out : (double*)malloc(sizeof(double) * np);
  memset(I, 0, sizeof(double) * np);
  for (int k = 0; k < nk; ++k) {
    for (int i = 0; i < T; ++i) {
      double ar = kr[(size_t)k * T + i], ai = ki[(size_t)k * T + i];
      A[2 * i] = ar * band[2 * i] - ai * band[2 * i + 1];
      A[2 * i + 1] = ar * band[2 * i + 1] + ai * band[2 * i];
    }
    double* Ek = E + 2 * np * k;
    synth_field(A, Nx, Ny, Lx, Ly, twx, twy, tmp, Ek);
    for (size_t p = 0; p < np; ++p)
      I[p] += (double)scales[k] * (Ek[2 * p] * Ek[2 * p] + Ek[2 * p + 1] * Ek[2 * p + 1]);
  }
  double* G = (double*)malloc(sizeof(double) * np);
  double L = resist_loss(I, tgt, np, a, thr, G);
  if (grad) {
    int cox = (Lx + 1) / 2, coy = (Ly + 1) / 2, cix = Lx / 2, ciy = Ly / 2;
    double* Gam = (double*)calloc(2 * T, sizeof(double));
    double* F = (double*)malloc(sizeof(double) * 2 * (size_t)Ly * Wx);
    for (int k = 0; k < nk; ++k) {
      double* Ek = E + 2 * np * k;
      /* F[y][nx] = sum_x G conj(E) e^{+i nx (x-cox)} */
      for (int y = 0; y < Ly; ++y)
        for (int nx = -Nx; nx <= Nx; ++nx) {
          double sr = 0.0, si = 0.0;
          for (int x = 0; x < Lx; ++x) {
            size_t p = (size_t)y * Lx + x;
            double hr = G[p] * Ek[2 * p], hi = -G[p] * Ek[2 * p + 1];
            int r = imod((long long)nx * (x - cox), Lx);
            sr += hr * twx[2 * r] - hi * twx[2 * r + 1];
            si += hr * twx[2 * r + 1] + hi * twx[2 * r];
          }
          F[2 * ((size_t)y * Wx + nx + Nx)] = sr;
          F[2 * ((size_t)y * Wx + nx + Nx) + 1] = si;
        }
      for (int ny = -Ny; ny <= Ny; ++ny)
        for (int nx = -Nx; nx <= Nx; ++nx) {
          double sr = 0.0, si = 0.0;
          for (int y = 0; y < Ly; ++y) {
            int r = imod((long long)ny * (y - coy), Ly);
            const double* f = F + 2 * ((size_t)y * Wx + nx + Nx);
            sr += f[0] * twy[2 * r] - f[1] * twy[2 * r + 1];
            si += f[0] * twy[2 * r + 1] + f[1] * twy[2 * r];
          }
          int i = (ny + Ny) * Wx + nx + Nx;
          double kre = kr[(size_t)k * T + i], kim = ki[(size_t)k * T + i];
          double l = (double)scales[k];
          Gam[2 * i] += l * (kre * sr - kim * si);
          Gam[2 * i + 1] += l * (kre * si + kim * sr);
        }
    }
    /* grad(x') = (2/(LxLy)) Re sum_n Gam[n] e^{-i n.(x'-c_in)} */
    double* H = (double*)malloc(sizeof(double) * 2 * (size_t)Wy * Lx);
    for (int ny = 0; ny < Wy; ++ny)
      for (int x = 0; x < Lx; ++x) {
        double sr = 0.0, si = 0.0;
        for (int nx = -Nx; nx <= Nx; ++nx) {
          int r = imod(-(long long)nx * (x - cix), Lx);
          const double* g = Gam + 2 * (ny * Wx + nx + Nx);
          sr += g[0] * twx[2 * r] - g[1] * twx[2 * r + 1];
          si += g[0] * twx[2 * r + 1] + g[1] * twx[2 * r];
        }
        H[2 * ((size_t)ny * Lx + x)] = sr;
        H[2 * ((size_t)ny * Lx + x) + 1] = si;
      }
    double sc = 2.0 / ((double)Lx * (double)Ly);
    for (int y = 0; y < Ly; ++y)
      for (int x = 0; x < Lx; ++x) {
        double acc = 0.0;
        for (int ny = -Ny; ny <= Ny; ++ny) {
          int r = imod(-(long long)ny * (y - ciy), Ly);
          const double* h = H + 2 * ((size_t)(ny + Ny) * Lx + x);
          acc += h[0] * twy[2 * r] - h[1] * twy[2 * r + 1];
        }
        grad[(size_t)y * Lx + x] = sc * acc;
      }
    free(Gam); free(F); free(H);
  }
  free(band); free(twx); free(twy); free(A); free(tmp); free(E); free(G);
  if (!img_out) free(I);
  return L;
}

/* Abbe image and source gradient.  For every source grid point (p,q) with
 * p^2+q^2 <= sh^2 the coherent field is
 *   E_s(x) = sum_n P(ID(-n), s) M^[n] e^{+2 pi i n.(x-c_out)/L},
 * (the reference stores kernels index-reversed, klitho_tcc.cpp:358-362, so
 * the SOCS image with every kernel equals sum_s S_s |E_s|^2).  With the
 * resist loss L of I_A = sum_s S_s |E_s|^2, dL/dS_s = sum_x G(x) |E_s(x)|^2,
 * evaluated here through the band form
 *   sum_x G |E_s|^2 = Lx Ly sum_{n,n'} A_s[n] conj(A_s[n']) conj(G^[n-n'])
 * with G^[g] = 1/(LxLy) sum_x G(x) e^{-2 pi i g.(x-c_out)/L}.
 * grid_grad: srcSize^2 (zero outside the disc).  Returns L. */
double or_abbe_loss_grad(const double* mask, const float* tgt, int Lx, int Ly,
                         const float* src, int srcSize, float lambda, float NA,
                         float defocus, int Nx, int Ny, float nmed, double a, double thr,
                         double* img_out, double* grid_grad) {
  pupil_geom g;
  geom_init(&g, srcSize, 1.0f, lambda, NA, defocus, Lx, Ly, nmed);
  int Wx = 2 * Nx + 1, Wy = 2 * Ny + 1, T = Wx * Wy;
  size_t np = (size_t)Lx * Ly;
  double* band = (double*)malloc(sizeof(double) * 2 * T);
  band_dft(NULL, mask, Lx, Ly, Nx, Ny, band);
  int Gx = 4 * Nx + 1, Gy = 4 * Ny + 1;
  double* J = (double*)calloc((size_t)Gx * Gy * 2, sizeof(double));
  double* As = (double*)malloc(sizeof(double) * 2 * T);
  int* nzl = (int*)malloc(sizeof(int) * T);
  /* pass 1: spectrum of I_A over the used source points */
  for (int q = -g.sh; q <= g.sh; ++q)
    for (int p = -g.sh; p <= g.sh; ++p) {
      if (p * p + q * q > g.sh * g.sh) continue;
      double S = src[(q + g.sh) * srcSize + p + g.sh];
      if (S == 0.0) continue;
      for (int ny = -Ny; ny <= Ny; ++ny)
        for (int nx = -Nx; nx <= Nx; ++nx) {
          double re = 0, im = 0;
          int i = (ny + Ny) * Wx + nx + Nx;
          if (pupil_at(&g, p, q, -nx, -ny, &re, &im)) {
            As[2 * i] = re * band[2 * i] - im * band[2 * i + 1];
            As[2 * i + 1] = re * band[2 * i + 1] + im * band[2 * i];
          } else {
            As[2 * i] = As[2 * i + 1] = 0.0;
          }
        }
      int nnz = 0;
      for (int i = 0; i < T; ++i)
        if (As[2 * i] != 0.0 || As[2 * i + 1] != 0.0) nzl[nnz++] = i;
      for (int u1 = 0; u1 < nnz; ++u1) {
        int n1 = nzl[u1];
        double ar = As[2 * n1], ai = As[2 * n1 + 1];
        int n1y = n1 / Wx - Ny, n1x = n1 % Wx - Nx;
        for (int u2 = 0; u2 < nnz; ++u2) {
          int n2 = nzl[u2];
          double br = As[2 * n2], bi = As[2 * n2 + 1];
          int n2y = n2 / Wx - Ny, n2x = n2 % Wx - Nx;
          int gx = n1x - n2x + 2 * Nx, gy = n1y - n2y + 2 * Ny;
          double* jj = J + 2 * ((size_t)gy * Gx + gx);
          jj[0] += S * (ar * br + ai * bi);
          jj[1] += S * (ai * br - ar * bi);
        }
      }
    }
  /* synthesise I_A at every pixel */
  double* twx = (double*)malloc(sizeof(double) * 2 * Lx);
  double* twy = (double*)malloc(sizeof(double) * 2 * Ly);
  twiddles(Lx, +1, twx);
  twiddles(Ly, +1, twy);
  double* I = img_out ? img_out : (double*)malloc(sizeof(double) * np);
  {
    double* tmp = (double*)malloc(sizeof(double) * 2 * (size_t)Gy * Lx);
    double* E = (double*)malloc(sizeof(double) * 2 * np);
    synth_field(J, 2 * Nx, 2 * Ny, Lx, Ly, twx, twy, tmp, E);
    for (size_t p = 0; p < np; ++p) I[p] = E[2 * p];
    free(tmp);
    free(E);
  }
  double* G = (double*)malloc(sizeof(double) * np);
  double L = resist_loss(I, tgt, np, a, thr, G);
  if (grid_grad) {
    /* G^ band over |g| <= 2N, centre c_out */
    double* Gh = (double*)malloc(sizeof(double) * 2 * (size_t)Gx * Gy);
    {
      int cx = (Lx + 1) / 2, cy = (Ly + 1) / 2;
      double* R = (double*)calloc((size_t)Ly * Gx * 2, sizeof(double));
      for (int y = 0; y < Ly; ++y)
        for (int x = 0; x < Lx; ++x) {
          double v = G[(size_t)y * Lx + x];
          for (int gx = -2 * Nx; gx <= 2 * Nx; ++gx) {
            int r = imod(-(long long)gx * (x - cx), Lx);
            R[2 * ((size_t)y * Gx + gx + 2 * Nx)] += v * twx[2 * r];
            R[2 * ((size_t)y * Gx + gx + 2 * Nx) + 1] += v * twx[2 * r + 1];
          }
        }
      double inv = 1.0 / ((double)Lx * Ly);
      for (int gy = -2 * Ny; gy <= 2 * Ny; ++gy)
        for (int gx = 0; gx < Gx; ++gx) {
          double sr = 0.0, si = 0.0;
          for (int y = 0; y < Ly; ++y) {
            int r = imod(-(long long)gy * (y - cy), Ly);
            const double* rr = R + 2 * ((size_t)y * Gx + gx);
            sr += rr[0] * twy[2 * r] - rr[1] * twy[2 * r + 1];
            si += rr[0] * twy[2 * r + 1] + rr[1] * twy[2 * r];
          }
          Gh[2 * ((size_t)(gy + 2 * Ny) * Gx + gx)] = sr * inv;
          Gh[2 * ((size_t)(gy + 2 * Ny) * Gx + gx) + 1] = si * inv;
        }
      free(R);
    }
    memset(grid_grad, 0, sizeof(double) * srcSize * srcSize);
    double LL = (double)Lx * Ly;
    for (int q = -g.sh; q <= g.sh; ++q)
      for (int p = -g.sh; p <= g.sh; ++p) {
        if (p * p + q * q > g.sh * g.sh) continue;
        for (int ny = -Ny; ny <= Ny; ++ny)
          for (int nx = -Nx; nx <= Nx; ++nx) {
            double re = 0, im = 0;
            int i = (ny + Ny) * Wx + nx + Nx;
            if (pupil_at(&g, p, q, -nx, -ny, &re, &im)) {
              As[2 * i] = re * band[2 * i] - im * band[2 * i + 1];
              As[2 * i + 1] = re * band[2 * i + 1] + im * band[2 * i];
            } else {
              As[2 * i] = As[2 * i + 1] = 0.0;
            }
          }
        double acc = 0.0;
        int nnz = 0;
        for (int i = 0; i < T; ++i)
          if (As[2 * i] != 0.0 || As[2 * i + 1] != 0.0) nzl[nnz++] = i;
        for (int u1 = 0; u1 < nnz; ++u1) {
          int n1 = nzl[u1];
          double ar = As[2 * n1], ai = As[2 * n1 + 1];
          int n1y = n1 / Wx - Ny, n1x = n1 % Wx - Nx;
          for (int u2 = 0; u2 < nnz; ++u2) {
            int n2 = nzl[u2];
            double br = As[2 * n2], bi = As[2 * n2 + 1];
            int n2y = n2 / Wx - Ny, n2x = n2 % Wx - Nx;
            /* A1 conj(A2) conj(G^[n1-n2]) */
            double pr = ar * br + ai * bi, pi = ai * br - ar * bi;
            const double* gg = Gh + 2 * ((size_t)(n1y - n2y + 2 * Ny) * Gx + n1x - n2x + 2 * Nx);
            acc += pr * gg[0] + pi * gg[1];
          }
        }
        grid_grad[(q + g.sh) * srcSize + p + g.sh] = LL * acc;
      }
    free(Gh);
  }
  free(band); free(J); free(As); free(nzl); free(twx); free(twy); free(G);
  if (!img_out) free(I);
  return L;
}

/* Abbe image only (no loss): I_A(x) = sum_s S_s |E_s(x)|^2 */
void or_abbe_image(const double* mask, int Lx, int Ly, const float* src, int srcSize,
                   float lambda, float NA, float defocus, int Nx, int Ny, float nmed,
                   double* img) {
  size_t np = (size_t)Lx * Ly;
  float* zt = (float*)calloc(np, sizeof(float));
  or_abbe_loss_grad(mask, zt, Lx, Ly, src, srcSize, lambda, NA, defocus, Nx, Ny, nmed,
                    1.0, 0.0, img, NULL);
  free(zt);
}

/* ------------------------------------------------------------------------ */
/* Resist step: the 7-level view image_s.png (common/file_io.cpp:25-80,      */
/* klitho_tcc.cpp:547 with min 0, max TARGET_INTENSITY = 0.225, :442)         */
/* ------------------------------------------------------------------------ */

/* lev[i] = the colour class 0..6 the reference gives img[i]: v =
 * std::max(0.0f, std::min(1.0f, (img - mn) / (mx - mn))) in float, classes
 * [k/6, (k+1)/6) compared in double, 6 for v >= 1; mn == mx gives class 3
 * (the reference's uniform green). */
void or_resist_levels(const float* img, int n, float mn, float mx, unsigned char* lev) {
  for (int i = 0; i < n; ++i) {
    if (mn == mx) {
      lev[i] = 3;
      continue;
    }
    float q = (img[i] - mn) / (mx - mn);
    float t = (q < 1.0f) ? q : 1.0f;
    double v = (0.0f < t) ? t : 0.0f;
    int c = 6;
    for (int k = 1; k <= 6; ++k)
      if (v < (double)k / 6.0) {
        c = k - 1;
        break;
      }
    lev[i] = (unsigned char)c;
  }
}
