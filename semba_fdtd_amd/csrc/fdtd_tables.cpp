// fdtd_tables.cpp -- host-side coefficient tables of the time-stepping path,
// built once and uploaded (never recomputed on the device).  Each function
// evaluates the reference's expressions with the reference's kinds: real(4)
// storage, real(8) dt, complex(8) pole coefficients, so the uploaded tables
// are the ones the Fortran code would hold.
#include <cmath>
#include <cstdint>
#include "sembafdtd.h"

namespace {

// calc_constants.F90:132-144 (and the GM twin): lossy update pair with the
// exponential fallback when the centred coefficient goes negative.
void lossy_pair(double dt, float s, float perm, float &c1, float &c2) {
  const double half = (double)s * dt / (double)(2.0f * perm);
  c1 = (float)((1.0 - half) / (1.0 + half));
  c2 = (float)(dt / (double)perm / (1.0 + half));
  if (c1 < 0.0f) {
    c1 = (float)std::exp(-(double)s * dt / (double)perm);
    c2 = (1.0f - c1) / s;
  }
}

// complex division with range reduction (Smith), as gfortran lowers it
struct cplx { double re, im; };
cplx cdiv(cplx a, cplx b) {
  if (std::fabs(b.re) >= std::fabs(b.im)) {
    const double r = b.im / b.re, d = b.re + b.im * r;
    return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  }
  const double r = b.re / b.im, d = b.re * r + b.im;
  return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}

// sigma_max of one CPML face, borderscpml.F90:1487-1506
float cpml_sigma_max(int32_t L, float m, float R, float del, float eps0, float mu0) {
  if (L == 0) return 1.0e29f;
  const float z0 = std::sqrt(mu0 / eps0);
  if (L == 5 || L == 10) return 0.8f * (m + 1.0f) / (z0 * del);
  return -((std::log(R) * (m + 1.0f)) / (2.0f * z0 * (float)L * del));
}

struct Face {
  int32_t L;
  float m, smax, amax, kmax;
};

// one profile point at normalised depth d (and 1-d for alpha):
// borderscpml.F90:1544-1556 (E) / 1626-1638 (H)
void profile_point(const Face &f, float d, float dI, float alphaorden, double dt,
                   float eps0, float delta, float &b, float &c, float &inv) {
  float s, k;
  if (f.m == 0.0f) {
    s = f.smax;
    k = 1.0f + (f.kmax - 1.0f);
  } else {
    const float dm = std::pow(d, f.m);
    s = f.smax * dm;
    k = 1.0f + (f.kmax - 1.0f) * std::pow(d, f.m);
  }
  const float a = f.amax * std::pow(dI, alphaorden);
  b = (float)std::exp((double)(-(s / k + a)) * dt / (double)eps0);
  c = (s * (b - 1.0f) / (s + k * a) / k) / delta;
  inv = 1.0f / (k * delta);
}

}  // namespace

extern "C" {

double fdtd_courant_dt(float dxmin, float dymin, float dzmin, float eps0, float mu0,
                       float cfl, int32_t forcecfl, double dt_in) {
  // semba_fdtd.F90:154-156, 1211-1250
  const float c0 = 1.0f / std::sqrt(eps0 * mu0);
  const float a = 1.0f / dxmin, b = 1.0f / dymin, c = 1.0f / dzmin;
  const float dtlay = 1.0f / (c0 * std::sqrt(a * a + b * b + c * c));
  if (forcecfl) return (double)(dtlay * cfl);
  const float lim = dtlay * 0.8f;
  return dt_in > (double)lim ? (double)lim : dt_in;
}

int fdtd_build_g_tables(int32_t nmedia, const float *epr, const float *sigma,
                        const float *mur, const float *sigmam, const int32_t *kind,
                        double dt, float eps0, float mu0, float *g1, float *g2,
                        float *gm1, float *gm2) {
  if (nmedia < 0 || !epr || !sigma || !mur || !sigmam || !kind) return FDTD_ERR_ARG;
  for (int32_t r = 0; r <= nmedia; ++r) {
    if (r == 0 || (kind[r] & FDTD_KIND_PEC)) {  // calc_constants.F90:62-64
      g1[r] = g2[r] = gm1[r] = gm2[r] = 0.0f;
      continue;
    }
    // one dispersive kind alone: placeholders the dispersive init
    // overwrites (:109-128); both kinds fall to the lossy branch (:129-140)
    const bool ed = kind[r] & FDTD_KIND_EDISP, md = kind[r] & FDTD_KIND_MDISP;
    if (ed && !md) {
      g1[r] = 0.0f;
      g2[r] = 0.0f;
    } else {
      lossy_pair(dt, sigma[r], eps0 * epr[r], g1[r], g2[r]);
    }
    if (md && !ed) {
      gm1[r] = 0.0f;
      gm2[r] = 0.0f;
    } else {
      lossy_pair(dt, sigmam[r], mu0 * mur[r], gm1[r], gm2[r]);
    }
  }
  return FDTD_OK;
}

int fdtd_build_spacing(int32_t n, const float *D, float *de, float *dh, float *ide,
                       float *idh) {
  // timestepping.F90:314-370: D covers [-2, n+1]; outputs cover [-1, n+1]
  if (n <= 0 || !D) return FDTD_ERR_ARG;
  for (int32_t i = -1; i <= n + 1; ++i) {
    const float prim = D[i + 2], prev = D[i + 1];
    de[i + 1] = prim;
    dh[i + 1] = (prim + prev) / 2.0f;
    ide[i + 1] = 1.0f / de[i + 1];
    idh[i + 1] = 1.0f / dh[i + 1];
  }
  return FDTD_OK;
}

int fdtd_build_lines(int32_t n, int32_t Llo, int32_t Lhi, const float *D, float *line) {
  // discretisation lines, preprocess_geom.F90:5330-5347 (interior: origin 0,
  // line(i+1) = d(i) + line(i)) and 5690-5712 (CPML and padding extended with
  // the first / last interior cell); D covers [-2, n+1], line [-1, n+1]
  if (n <= 0 || Llo < 0 || Lhi < 0 || Llo + Lhi >= n || !D || !line) return FDTD_ERR_ARG;
  const int32_t lo = Llo, hi = n - Lhi;
  line[lo + 1] = 0.0f;
  for (int32_t i = lo; i < hi; ++i) line[i + 2] = D[i + 2] + line[i + 1];
  float delta = line[lo + 2] - line[lo + 1];
  for (int32_t i = lo - 1; i >= -1; --i) line[i + 1] = line[i + 2] - delta;
  delta = line[hi + 1] - line[hi];
  for (int32_t i = hi + 1; i <= n + 1; ++i) line[i + 1] = line[i] + delta;
  return FDTD_OK;
}

int fdtd_build_cpml_axis(int32_t n, int32_t Llo, int32_t Lhi, float orden_lo,
                         float orden_hi, float R_lo, float R_hi, float alphamaxpar,
                         float alphaorden, float kappamaxpar, double dt, float eps0,
                         float mu0, const float *de, const float *dh, float *pbe,
                         float *pce, float *pbm, float *pcm, float *ide, float *idh) {
  if (n <= 0 || Llo < 0 || Lhi < 0 || Llo + Lhi > n) return FDTD_ERR_ARG;
  // borderscpml.F90:1465-1733 in total coordinates (interior [Llo, n-Lhi])
  Face face[2];
  const int32_t Ls[2] = {Llo, Lhi};
  const float ord[2] = {orden_lo, orden_hi}, Rs[2] = {R_lo, R_hi};
  const float del[2] = {de[0], de[n + 1]};  // dxe at Alloc(iEx) XI / XE
  for (int s = 0; s < 2; ++s) {
    face[s].L = Ls[s];
    face[s].m = ord[s];
    face[s].smax = cpml_sigma_max(Ls[s], ord[s], Rs[s], del[s], eps0, mu0);
    face[s].amax = Ls[s] ? alphamaxpar * face[s].smax : 0.0f;
    face[s].kmax = Ls[s] ? kappamaxpar : 1.0f;
  }
  for (int32_t i = 0; i < n + 3; ++i) {
    pbe[i] = 1.0f; pce[i] = 0.0f; pbm[i] = 1.0f; pcm[i] = 0.0f;
  }
  const int32_t lo = Llo, hi = n - Lhi;
  for (int32_t i = -1; i <= n; ++i) {  // E side, integer depth (:1544-1559)
    int s;
    int32_t depth;
    if (i <= lo - 1) { s = 0; depth = lo - i; }
    else if (i >= hi + 1) { s = 1; depth = i - hi; }
    else continue;
    const float L = (float)Ls[s];
    const float d = 1.0f * (float)depth / L;
    const float dI = 1.0f * (float)(Ls[s] - depth) / L;
    profile_point(face[s], d, dI, alphaorden, dt, eps0, dh[i + 1], pbe[i + 1],
                  pce[i + 1], idh[i + 1]);
  }
  for (int32_t i = -1; i <= n + 1; ++i) {  // H side, half-cell depth (:1626-1643)
    int s;
    float depth;
    if (i <= lo - 1) { s = 0; depth = (float)lo - ((float)i + 0.5f); }
    else if (i >= hi) { s = 1; depth = (float)(i - hi) + 0.5f; }
    else continue;
    const float L = (float)Ls[s];
    const float d = 1.0f * depth / L;
    const float dI = 1.0f * ((float)Ls[s] - depth) / L;
    profile_point(face[s], d, dI, alphaorden, dt, eps0, de[i + 1], pbm[i + 1],
                  pcm[i + 1], ide[i + 1]);
  }
  return FDTD_OK;
}

int fdtd_build_ade(int32_t npoles, const double *c, const double *a, float eps11,
                   float sigma11, double dt, double *kappa, double *beta, double *g3,
                   float *g1, float *g2) {
  // electricdispersive.F90:98-146
  if (npoles < 0 || npoles > 8) return FDTD_ERR_ARG;
  for (int32_t p = 0; p < npoles; ++p) {
    const cplx ah = {a[2 * p] * dt / 2.0, a[2 * p + 1] * dt / 2.0};
    const cplx den = {1.0 - ah.re, -ah.im};
    const cplx k = cdiv({1.0 + ah.re, ah.im}, den);
    const cplx b = cdiv({c[2 * p] * dt, c[2 * p + 1] * dt}, den);
    kappa[2 * p] = k.re; kappa[2 * p + 1] = k.im;
    beta[2 * p] = b.re; beta[2 * p + 1] = b.im;
  }
  float tempo = 0.0f;
  for (int32_t p = 0; p < npoles; ++p) tempo = (float)((double)tempo + beta[2 * p]);
  const float e2 = 2.0f * eps11 + tempo;
  const double sdt = (double)sigma11 * dt;
  *g1 = (float)(((double)e2 - sdt) / ((double)e2 + sdt));
  *g2 = (float)(2.0 * dt / ((double)e2 + sdt));
  const double half = (double)(*g2 / 2.0f);
  for (int32_t p = 0; p < npoles; ++p) {
    g3[2 * p] = half * (1.0 + kappa[2 * p]);
    g3[2 * p + 1] = half * kappa[2 * p + 1];
  }
  return FDTD_OK;
}

int fdtd_build_mur(int32_t nmedia, const float *epr, const float *mur, double dt, float eps0,
                   float mu0, const float idh_face[6], float *cab1) {
  // calc_murconstants, bordersmur.F90:316-363: per face and medium
  // cnum = (1/Idh at the face) / (dt * c0 / sqrt(epr*mur)), real(4) result
  if (nmedia < 0 || !epr || !mur || !idh_face || !cab1) return FDTD_ERR_ARG;
  const float c0 = 1.0f / std::sqrt(eps0 * mu0);
  for (int f = 0; f < 6; ++f)
    for (int32_t m = 0; m <= nmedia; ++m) {
      const float v = std::sqrt(epr[m] * mur[m]);
      const float cnum = (float)((double)(1.0f / idh_face[f]) / (dt * (double)c0 / (double)v));
      cab1[(size_t)f * (nmedia + 1) + m] = (1.0f - cnum) / (1.0f + cnum);
    }
  return FDTD_OK;
}

// evolucion, nodalsources.F90:317-356 (real(4) linear interpolation)
float fdtd_evolucion(float t, float deltaevol, int32_t numus, const float *evol) {
  const long long nprev = (long long)(t / deltaevol);
  if (nprev + 1 > numus || nprev + 1 <= 0) return 0.0f;
  return (evol[nprev + 1] - evol[nprev]) / deltaevol * (t - (float)nprev * deltaevol) +
         evol[nprev];
}

}  // extern "C"
