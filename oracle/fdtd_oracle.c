/*
 * fdtd_oracle.c -- CPU restatement of semba-fdtd's field time-stepping path
 * (Yee E/H curl updates, CPML, electric pole-residue ADE, nodal sources,
 * coefficient tables).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  Build it
 * with -ffp-contract=off and no fast-math so that every expression is
 * evaluated in the reference's order with one rounding per operation; the
 * reference's default x86-64 build has no FMA either.
 *
 * Every function names the reference lines it restates (paths relative to the
 * reference root).  The real kind is a build parameter, as in the reference
 * (-DCompileWithReal4 / -DCompileWithReal8, fdetypes.F90:140-162): the
 * default build keeps fields and tables in real(4) with time and dt in
 * real(8) and dispersive currents in complex(8); -DORC_REAL64 builds the
 * real(8) variant used as the FP64 yardstick for error statements.
 *
 * TF/SF plane waves follow planewaves.F90 (InitPlaneWave, Incid,
 * AdvancePlaneWaveE/H).
 *
 * Parity pins (see tests/): tables and a full step loop are compared
 * bit-for-bit against the reference's own Fortran modules built from source
 * (oracle/_ref), and against closed forms (Yee cavity eigenfrequencies,
 * CPML reflection, discrete pole-residue permittivity).
 */
#include "fdtd_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef orc_real real;
#define R(x) ((real)(x))
#ifdef ORC_REAL64
#define RSQRT sqrt
#define RLOG log
#define RPOW pow
#else
#define RSQRT sqrtf
#define RLOG logf
#define RPOW powf
#endif

#define AX(g) ((size_t)(g)->nx + 3)
#define AY(g) ((size_t)(g)->ny + 3)
#define IDX(g, i, j, k) \
  ((((size_t)((k) + 1)) * AY(g) + (size_t)((j) + 1)) * AX(g) + (size_t)((i) + 1))

/* media kind bits: PEC, PMC, electric dispersive, magnetic dispersive */
#define ORC_KIND_PEC 1
#define ORC_KIND_PMC 2
#define ORC_KIND_EDISP 4
#define ORC_KIND_MDISP 8

/* --------------------------------------------------------------------------
 * dt: semba_fdtd.F90:1211-1250.  dtlay is formed in the real kind from the
 * minimum cell sizes; with -cfl the step is dtlay*cfl (stored in the real(8)
 * sgg%dt), otherwise the input dt is clipped to heurCFL*dtlay (heurCFL = 0.8,
 * fdetypes.F90:186).  cluz = 1/sqrt(eps0*mu0) (semba_fdtd.F90:154-156).
 * ------------------------------------------------------------------------ */
double orc_dt(real dxmin, real dymin, real dzmin, real eps0, real mu0, real cfl,
              int32_t forcecfl, double dt_in) {
  real cluz = R(1.0) / RSQRT(eps0 * mu0);
  real ix = R(1.0) / dxmin, iy = R(1.0) / dymin, iz = R(1.0) / dzmin;
  real dtlay = R(1.0) / (cluz * RSQRT(ix * ix + iy * iy + iz * iz));
  if (forcecfl) return (double)(dtlay * cfl);
  if (dt_in > (double)(dtlay * R(0.8))) return (double)(dtlay * R(0.8));
  return dt_in;
}

/* --------------------------------------------------------------------------
 * Per-medium update coefficients: calc_constants.F90:36-148.
 * kind: 0 lossy dielectric (:132-144), ORC_KIND_PEC (or r == 0) all zero
 * (:62-64), ORC_KIND_EDISP alone: G1/G2 placeholders 0 (overwritten by the
 * ADE init), GM from mu/sigmam (:109-118); ORC_KIND_MDISP alone: the mirror
 * image (:119-128); both: lossy G and GM, both overwritten by the inits.  Expressions mix the real kind with
 * the real(8) dt exactly as written there.
 * ------------------------------------------------------------------------ */
static void orc_lossy(double dt, real sig, real eps, real *g1, real *g2) {
  /* G1=(1 - S*dt/(2*eps))/(1 + S*dt/(2*eps)); G2=dt/eps/(1 + S*dt/(2*eps)) */
  double a = (double)sig * dt / (double)(R(2.0) * eps);
  double x1 = (1.0 - a) / (1.0 + a);
  double x2 = dt / (double)eps / (1.0 + a);
  *g1 = (real)x1;
  *g2 = (real)x2;
  if (*g1 < R(0.0)) { /* exponential time stepping, :135-138 */
    *g1 = (real)exp(-(double)sig * dt / (double)eps);
    *g2 = (R(1.0) - *g1) / sig;
  }
}

void orc_g_tables(int32_t nmedia, const real *epr, const real *sigma, const real *mur,
                  const real *sigmam, const int32_t *kind, double dt, real eps0, real mu0,
                  real *g1, real *g2, real *gm1, real *gm2) {
  for (int32_t r = 0; r <= nmedia; ++r) {
    real eps = eps0 * epr[r];
    real mu = mu0 * mur[r];
    const int ed = (kind[r] & ORC_KIND_EDISP) != 0, md = (kind[r] & ORC_KIND_MDISP) != 0;
    if ((kind[r] & ORC_KIND_PEC) || r == 0) {
      g1[r] = g2[r] = gm1[r] = gm2[r] = R(0.0);
    } else if (ed && !md) {
      g1[r] = R(0.0);
      g2[r] = R(0.0);
      orc_lossy(dt, sigmam[r], mu, &gm1[r], &gm2[r]);
    } else if (md && !ed) { /* :119-128, GM placeholders for InitMdispersives */
      orc_lossy(dt, sigma[r], eps, &g1[r], &g2[r]);
      gm1[r] = R(0.0);
      gm2[r] = R(0.0);
    } else {
      orc_lossy(dt, sigma[r], eps, &g1[r], &g2[r]);
      orc_lossy(dt, sigmam[r], mu, &gm1[r], &gm2[r]);
    }
  }
}

/* --------------------------------------------------------------------------
 * Spacings: timestepping.F90:314-370.  dxe(i) = DX(i) (primal, H side),
 * dxh(i) = (DX(i) + DX(i-1))/2 (dual, E side), inverses, over the Alloc
 * range [-1, n+1].  DX is indexed [-2, n+1] (DX[i+2]).
 * ------------------------------------------------------------------------ */
void orc_spacing(int32_t n, const real *DX, real *de, real *dh, real *Ide, real *Idh) {
  for (int32_t i = -1; i <= n + 1; ++i) {
    de[i + 1] = DX[i + 2];
    dh[i + 1] = (DX[i + 2] + DX[i + 1]) / R(2.0);
    Ide[i + 1] = R(1.0) / de[i + 1];
    Idh[i + 1] = R(1.0) / dh[i + 1];
  }
}

/* --------------------------------------------------------------------------
 * CPML profiles along one axis.
 * Depths: borderscpml.F90:228-300 (integer depth ce for E, half-cell cm for
 * H).  sigma_max / alpha / kappa: calc_cpmlconstants, borderscpml.F90:
 * 1465-1733, including the 5/10-layer Gedney optimum (:1494-1497) and the
 * absent-side value 1e29 (:1504).  b = exp(-(s/k + a)*dt/eps0) in real(8)
 * stored in the real kind; c = s*(b-1)/(s + k*a)/k / delta; Id = 1/(k*delta).
 * E uses i <= SIN_lo-1 / i >= SIN_hi+1 (:1545,1557); H uses i <= SIN_lo-1 /
 * i >= SIN_hi (:1627,1643), with SIN_lo = Llo, SIN_hi = n - Lhi in total
 * coordinates.  Arrays have length n+3 (index i+1).
 * ------------------------------------------------------------------------ */
static real orc_sigmax(int32_t L, real orden, real R_, real del, real eps0, real mu0) {
  if (L == 0) return R(1.0e29);
  if (L == 10 || L == 5) return R(0.8) * (orden + R(1.0)) / (RSQRT(mu0 / eps0) * del);
  return -((RLOG(R_) * (orden + R(1.0))) / (R(2.0) * RSQRT(mu0 / eps0) * (real)L * del));
}

void orc_cpml_axis(int32_t n, int32_t Llo, int32_t Lhi, real orden_lo, real orden_hi,
                   real R_lo, real R_hi, real alphamaxpar, real alphaorden,
                   real kappamaxpar, double dt, real eps0, real mu0, const real *de,
                   const real *dh, real *pbe, real *pce, real *pbm, real *pcm, real *Ide,
                   real *Idh) {
  const int32_t SI = Llo, SE = n - Lhi;
  /* del: dxe at Alloc(iEx) XI / XE = -1 / n (:1480-1485) */
  real smax[2], amax[2], kmax[2];
  real dlo = de[-1 + 1], dhi = de[n + 1];
  smax[0] = orc_sigmax(Llo, orden_lo, R_lo, dlo, eps0, mu0);
  smax[1] = orc_sigmax(Lhi, orden_hi, R_hi, dhi, eps0, mu0);
  amax[0] = Llo ? alphamaxpar * smax[0] : R(0.0);
  amax[1] = Lhi ? alphamaxpar * smax[1] : R(0.0);
  kmax[0] = Llo ? kappamaxpar : R(1.0);
  kmax[1] = Lhi ? kappamaxpar : R(1.0);
  const real ord[2] = {orden_lo, orden_hi};
  const int32_t Ls[2] = {Llo, Lhi};

  for (int32_t i = -1; i <= n + 1; ++i) {
    pbe[i + 1] = R(1.0); pce[i + 1] = R(0.0);
    pbm[i + 1] = R(1.0); pcm[i + 1] = R(0.0);
  }
  /* electric: i over Alloc(iEx) = [-1, n] */
  for (int32_t i = -1; i <= n; ++i) {
    int side = -1;
    real ce = R(0.0), Ice = R(0.0);
    if (i <= SI - 1) {
      side = 0;
      ce = R(1.0) * (real)(SI - i) / (real)Ls[0];
      Ice = R(1.0) * (real)(Ls[0] - (SI - i)) / (real)Ls[0];
    } else if (i >= SE + 1) {
      side = 1;
      ce = R(1.0) * (real)(i - SE) / (real)Ls[1];
      Ice = R(1.0) * (real)(Ls[1] - (i - SE)) / (real)Ls[1];
    }
    if (side < 0) continue;
    real s, k;
    if (ord[side] == R(0.0)) {
      s = smax[side];
      k = R(1.0) + (kmax[side] - R(1.0));
    } else {
      s = smax[side] * RPOW(ce, ord[side]);
      k = R(1.0) + (kmax[side] - R(1.0)) * RPOW(ce, ord[side]);
    }
    real a = amax[side] * RPOW(Ice, alphaorden);
    real b = (real)exp((double)(-(s / k + a)) * dt / (double)eps0);
    pbe[i + 1] = b;
    pce[i + 1] = (s * (b - R(1.0)) / (s + k * a) / k) / dh[i + 1];
    Idh[i + 1] = R(1.0) / (k * dh[i + 1]);
  }
  /* magnetic: i over Alloc(iHx) = [-1, n+1] */
  for (int32_t i = -1; i <= n + 1; ++i) {
    int side = -1;
    real cm = R(0.0), Icm = R(0.0);
    if (i <= SI - 1) {
      side = 0;
      cm = R(1.0) * ((real)SI - ((real)i + R(0.5))) / (real)Ls[0];
      Icm = R(1.0) * ((real)Ls[0] - ((real)SI - ((real)i + R(0.5)))) / (real)Ls[0];
    } else if (i >= SE) {
      side = 1;
      cm = R(1.0) * ((real)(i - SE) + R(0.5)) / (real)Ls[1];
      Icm = R(1.0) * ((real)Ls[1] - ((real)(i - SE) + R(0.5))) / (real)Ls[1];
    }
    if (side < 0) continue;
    real s, k;
    if (ord[side] == R(0.0)) {
      s = smax[side];
      k = R(1.0) + (kmax[side] - R(1.0));
    } else {
      s = smax[side] * RPOW(cm, ord[side]);
      k = R(1.0) + (kmax[side] - R(1.0)) * RPOW(cm, ord[side]);
    }
    real a = amax[side] * RPOW(Icm, alphaorden);
    real b = (real)exp((double)(-(s / k + a)) * dt / (double)eps0);
    pbm[i + 1] = b;
    pcm[i + 1] = (s * (b - R(1.0)) / (s + k * a) / k) / de[i + 1];
    Ide[i + 1] = R(1.0) / (k * de[i + 1]);
  }
}

/* --------------------------------------------------------------------------
 * Pole-residue ADE coefficients: electricdispersive.F90:98-146.
 * kappa_p = (1 + a_p dt/2)/(1 - a_p dt/2), beta_p = c_p dt/(1 - a_p dt/2)
 * (complex(8)); tempo = sum Re(beta_p) accumulated in the real kind;
 * G1 = (2 eps + tempo - s dt)/(2 eps + tempo + s dt),
 * G2 = 2 dt/(2 eps + tempo + s dt)   (real(8) expressions stored real);
 * G3_p = G2/2 * (1 + kappa_p) with G2/2 in the real kind.
 * Complex division uses Smith's range-reduced form (gfortran's
 * -fcx-fortran-rules lowering).  Arrays are [p][2] (re, im).
 * ------------------------------------------------------------------------ */
static void cdiv(double a, double b, double c, double d, double *re, double *im) {
  if (fabs(c) >= fabs(d)) {
    double r = d / c, den = c + d * r;
    *re = (a + b * r) / den;
    *im = (b - a * r) / den;
  } else {
    double r = c / d, den = c * r + d;
    *re = (a * r + b) / den;
    *im = (b * r - a) / den;
  }
}

void orc_ade_tables(int32_t npoles, const double *c, const double *a, real eps11,
                    real sigma11, double dt, double *kappa, double *beta, double *g3,
                    real *g1_out, real *g2_out) {
  for (int32_t p = 0; p < npoles; ++p) {
    double ar = a[2 * p], ai = a[2 * p + 1];
    /* a*dt/2 : complex * real(8), then / 2.0 */
    double hr = ar * dt / 2.0, hi = ai * dt / 2.0;
    double nr = 1.0 + hr, ni = hi;
    double dr = 1.0 - hr, di = -hi;
    cdiv(nr, ni, dr, di, &kappa[2 * p], &kappa[2 * p + 1]);
    double cr = c[2 * p] * dt, ci = c[2 * p + 1] * dt;
    cdiv(cr, ci, dr, di, &beta[2 * p], &beta[2 * p + 1]);
  }
  real tempo = R(0.0);
  for (int32_t p = 0; p < npoles; ++p) tempo = (real)((double)tempo + beta[2 * p]);
  double sdt = (double)sigma11 * dt;
  real e2 = R(2.0) * eps11 + tempo;
  real g1 = (real)(((double)e2 - sdt) / ((double)e2 + sdt));
  real g2 = (real)(2.0 * dt / ((double)e2 + sdt));
  *g1_out = g1;
  *g2_out = g2;
  real h = g2 / R(2.0);
  for (int32_t p = 0; p < npoles; ++p) {
    double kr = 1.0 + kappa[2 * p], ki = kappa[2 * p + 1];
    g3[2 * p] = (double)h * kr;
    g3[2 * p + 1] = (double)h * ki;
  }
}

/* crea_timevector, timestepping.F90:3230-3240: tiempo(0)=0, tiempo(i) =
 * tiempo(i-1)+dt accumulated in real(8). */
void orc_time_vector(int32_t nlast, double dt, double *tiempo) {
  tiempo[0] = 0.0;
  for (int32_t i = 1; i <= nlast; ++i) tiempo[i] = tiempo[i - 1] + dt;
}

/* --------------------------------------------------------------------------
 * Yee E update: Advance_Ex/Ey/Ez, timestepping.F90:2165-2285.
 * Sweeps (total coordinates): Ex x[0,nx-1] y[0,ny] z[0,nz]; Ey x[0,nx]
 * y[0,ny-1] z[0,nz]; Ez x[0,nx] y[0,ny] z[0,nz-1].  OpenMP over (k, j) as
 * the reference's collapse(2) loops.
 * ------------------------------------------------------------------------ */
void orc_yee_e(orc_grid *g) {
  real *Ex = g->f[0], *Ey = g->f[1], *Ez = g->f[2];
  const real *Hx = g->f[3], *Hy = g->f[4], *Hz = g->f[5];
  const real *Idxh = g->id[3], *Idyh = g->id[4], *Idzh = g->id[5];
  const size_t sx = 1, sy = AX(g), sz = AX(g) * AY(g);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k <= g->nz; ++k)
    for (int j = 0; j <= g->ny; ++j)
      for (int i = 0; i <= g->nx - 1; ++i) {
        size_t p = IDX(g, i, j, k);
        int m = g->mi[0][p];
        Ex[p] = g->g1[m] * Ex[p] + g->g2[m] * ((Hz[p] - Hz[p - sy]) * Idyh[j + 1] -
                                              (Hy[p] - Hy[p - sz]) * Idzh[k + 1]);
      }
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k <= g->nz; ++k)
    for (int j = 0; j <= g->ny - 1; ++j)
      for (int i = 0; i <= g->nx; ++i) {
        size_t p = IDX(g, i, j, k);
        int m = g->mi[1][p];
        Ey[p] = g->g1[m] * Ey[p] + g->g2[m] * ((Hx[p] - Hx[p - sz]) * Idzh[k + 1] -
                                              (Hz[p] - Hz[p - sx]) * Idxh[i + 1]);
      }
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k <= g->nz - 1; ++k)
    for (int j = 0; j <= g->ny; ++j)
      for (int i = 0; i <= g->nx; ++i) {
        size_t p = IDX(g, i, j, k);
        int m = g->mi[2][p];
        Ez[p] = g->g1[m] * Ez[p] + g->g2[m] * ((Hy[p] - Hy[p - sx]) * Idxh[i + 1] -
                                              (Hx[p] - Hx[p - sy]) * Idyh[j + 1]);
      }
}

/* Yee H update: Advance_Hx/Hy/Hz, timestepping.F90:2294-2409.
 * Sweeps: Hx x[0,nx] y[0,ny-1] z[0,nz-1]; Hy x[0,nx-1] y[0,ny] z[0,nz-1];
 * Hz x[0,nx-1] y[0,ny-1] z[0,nz]. */
void orc_yee_h(orc_grid *g) {
  const real *Ex = g->f[0], *Ey = g->f[1], *Ez = g->f[2];
  real *Hx = g->f[3], *Hy = g->f[4], *Hz = g->f[5];
  const real *Idxe = g->id[0], *Idye = g->id[1], *Idze = g->id[2];
  const size_t sx = 1, sy = AX(g), sz = AX(g) * AY(g);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k <= g->nz - 1; ++k)
    for (int j = 0; j <= g->ny - 1; ++j)
      for (int i = 0; i <= g->nx; ++i) {
        size_t p = IDX(g, i, j, k);
        int m = g->mi[3][p];
        Hx[p] = g->gm1[m] * Hx[p] + g->gm2[m] * ((Ey[p + sz] - Ey[p]) * Idze[k + 1] -
                                                (Ez[p + sy] - Ez[p]) * Idye[j + 1]);
      }
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k <= g->nz - 1; ++k)
    for (int j = 0; j <= g->ny; ++j)
      for (int i = 0; i <= g->nx - 1; ++i) {
        size_t p = IDX(g, i, j, k);
        int m = g->mi[4][p];
        Hy[p] = g->gm1[m] * Hy[p] + g->gm2[m] * ((Ez[p + sx] - Ez[p]) * Idxe[i + 1] -
                                                (Ex[p + sz] - Ex[p]) * Idze[k + 1]);
      }
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k <= g->nz; ++k)
    for (int j = 0; j <= g->ny - 1; ++j)
      for (int i = 0; i <= g->nx - 1; ++i) {
        size_t p = IDX(g, i, j, k);
        int m = g->mi[5][p];
        Hz[p] = g->gm1[m] * Hz[p] + g->gm2[m] * ((Ex[p + sy] - Ex[p]) * Idye[j + 1] -
                                                (Ey[p + sx] - Ey[p]) * Idxe[i + 1]);
      }
}

/* --------------------------------------------------------------------------
 * CPML corrections.  Region boxes PMLc: borderscpml.F90:149-193 (each slab
 * spans the full sweep of the other two axes, so edges/corners overlap);
 * updates AdvanceelectricCPML :609-871 / AdvanceMagneticCPML :907-1171.
 * Region order left,right (y) ; down,up (z) ; back,front (x).
 * ------------------------------------------------------------------------ */
typedef struct { int lo[3], hi[3]; } box3;

/* sweep of component c (0..5) along axis a */
static int sweep_hi(const orc_grid *g, int c, int a) {
  int n = a == 0 ? g->nx : (a == 1 ? g->ny : g->nz);
  int node = (c < 3) ? (c != a) : (c - 3 == a);
  return node ? n : n - 1;
}

/* slab box of component c on axis a, side s (0 low, 1 high); 0 if empty */
static int slab_box(const orc_grid *g, int c, int a, int s, box3 *b) {
  if (g->L[a][s] == 0) return 0;
  for (int d = 0; d < 3; ++d) { b->lo[d] = 0; b->hi[d] = sweep_hi(g, c, d); }
  int n = a == 0 ? g->nx : (a == 1 ? g->ny : g->nz);
  int node = (c < 3) ? (c != a) : (c - 3 == a);
  int sin_hi = node ? n - g->L[a][1] : n - g->L[a][1] - 1; /* SINPML end */
  if (s == 0) {
    int e = g->L[a][0] - 1; /* SINPML start - 1 */
    if (e < b->hi[a]) b->hi[a] = e;
  } else {
    int st = sin_hi + 1;
    if (st > b->lo[a]) b->lo[a] = st;
  }
  return b->lo[a] <= b->hi[a];
}

/* one E (or H) psi loop over a slab; sign applies to the field correction.
 * src: field differenced along axis a; forward: H-side difference
 * src(p+s)-src(p), otherwise src(p)-src(p-s). */
static void cpml_loop(orc_grid *g, int c, int a, int side, int psislot, int src,
                      int forward, int sign) {
  box3 b;
  if (!slab_box(g, c, a, side, &b)) return;
  real *F = g->f[c];
  const real *S = g->f[src];
  real *psi = g->psi[psislot];
  const int16_t *mi = g->mi[c];
  const real *gg = (c < 3) ? g->g2 : g->gm2;
  const real *pb = (c < 3) ? g->pbe[a] : g->pbm[a];
  const real *pc = (c < 3) ? g->pce[a] : g->pcm[a];
  const size_t st[3] = {1, AX(g), AX(g) * AY(g)};
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = b.lo[2]; k <= b.hi[2]; ++k)
    for (int j = b.lo[1]; j <= b.hi[1]; ++j)
      for (int i = b.lo[0]; i <= b.hi[0]; ++i) {
        size_t p = IDX(g, i, j, k);
        int q = a == 0 ? i : (a == 1 ? j : k);
        real d = forward ? (S[p + st[a]] - S[p]) : (S[p] - S[p - st[a]]);
        psi[p] = pb[q + 1] * psi[p] + d * pc[q + 1];
        int m = mi[p];
        if (sign > 0)
          F[p] = F[p] + gg[m] * psi[p];
        else
          F[p] = F[p] - gg[m] * psi[p];
      }
}

void orc_cpml_e(orc_grid *g) {
  for (int s = 0; s < 2; ++s) { /* left, right */
    cpml_loop(g, ORC_EX, 1, s, ORC_PSI_EXY, ORC_HZ, 0, +1);
    cpml_loop(g, ORC_EZ, 1, s, ORC_PSI_EZY, ORC_HX, 0, -1);
  }
  for (int s = 0; s < 2; ++s) { /* down, up */
    cpml_loop(g, ORC_EY, 2, s, ORC_PSI_EYZ, ORC_HX, 0, +1);
    cpml_loop(g, ORC_EX, 2, s, ORC_PSI_EXZ, ORC_HY, 0, -1);
  }
  for (int s = 0; s < 2; ++s) { /* back, front */
    cpml_loop(g, ORC_EZ, 0, s, ORC_PSI_EZX, ORC_HY, 0, +1);
    cpml_loop(g, ORC_EY, 0, s, ORC_PSI_EYX, ORC_HZ, 0, -1);
  }
}

void orc_cpml_h(orc_grid *g) {
  for (int s = 0; s < 2; ++s) { /* left, right */
    cpml_loop(g, ORC_HX, 1, s, ORC_PSI_HXY, ORC_EZ, 1, -1);
    cpml_loop(g, ORC_HZ, 1, s, ORC_PSI_HZY, ORC_EX, 1, +1);
  }
  for (int s = 0; s < 2; ++s) { /* down, up */
    cpml_loop(g, ORC_HY, 2, s, ORC_PSI_HYZ, ORC_EX, 1, -1);
    cpml_loop(g, ORC_HX, 2, s, ORC_PSI_HXZ, ORC_EY, 1, +1);
  }
  for (int s = 0; s < 2; ++s) { /* back, front */
    cpml_loop(g, ORC_HZ, 0, s, ORC_PSI_HZX, ORC_EY, 1, -1);
    cpml_loop(g, ORC_HY, 0, s, ORC_PSI_HYX, ORC_EZ, 1, +1);
  }
}

/* --------------------------------------------------------------------------
 * ADE: AdvanceEDispersiveE, electricdispersive.F90:304-362, and
 * AdvanceMdispersiveH, magneticdispersive.F90:306-362 (the same statements on
 * H with GM3).  Per dispersive edge: F <- F - Re(G3_p J_p) one pole at a time
 * (real(8) product, rounded to the real kind after each pole); then J_p <-
 * kappa_p J_p + beta_p (F - Fprev)/dt with (F - Fprev) in the real kind; then
 * Fprev <- F.  Edges are those of the field's sweeps whose media index equals
 * the dispersive medium (electricdispersive.F90:148-244,
 * magneticdispersive.F90:139-265); the reference walks them serially.
 * ------------------------------------------------------------------------ */
static void ade_field(orc_grid *g, int side) {
  const int nd = side ? g->nmdisp : g->ndisp;
  for (int d = 0; d < nd; ++d) {
    const int med = side ? g->mdisp_medium[d] : g->disp_medium[d];
    const int np = side ? g->mdisp_npoles[d] : g->disp_npoles[d];
    const size_t o = (size_t)d * ORC_PMAX * 2;
    const double *K = (side ? g->mdisp_kappa : g->disp_kappa) + o;
    const double *B = (side ? g->mdisp_beta : g->disp_beta) + o;
    const double *G = (side ? g->mdisp_gm3 : g->disp_g3) + o;
    for (int c = 0; c < 3; ++c) {
      const int f = 3 * side + c;
      real *F = g->f[f];
      real *Ep = side ? g->hprev[c] : g->eprev[c];
      double *J = side ? g->mjcur[c] : g->jcur[c];
      int xe = sweep_hi(g, f, 0), ye = sweep_hi(g, f, 1), ze = sweep_hi(g, f, 2);
      for (int k = 0; k <= ze; ++k)
        for (int j = 0; j <= ye; ++j)
          for (int i = 0; i <= xe; ++i) {
            size_t p = IDX(g, i, j, k);
            if (g->mi[f][p] != med) continue;
            double *Jp = J + p * (size_t)g->jstride * 2;
            real e = F[p];
            for (int q = 0; q < np; ++q) {
              double re = G[2 * q] * Jp[2 * q] - G[2 * q + 1] * Jp[2 * q + 1];
              e = (real)((double)e - re);
            }
            F[p] = e;
            real diff = e - Ep[p];
            for (int q = 0; q < np; ++q) {
              double kr = K[2 * q], ki = K[2 * q + 1];
              double jr = Jp[2 * q], ji = Jp[2 * q + 1];
              double ar = kr * jr - ki * ji;
              double ai = kr * ji + ki * jr;
              double br = B[2 * q] * (double)diff / g->dt;
              double bi = B[2 * q + 1] * (double)diff / g->dt;
              Jp[2 * q] = ar + br;
              Jp[2 * q + 1] = ai + bi;
            }
            Ep[p] = e;
          }
    }
  }
}

void orc_ade_e(orc_grid *g) { ade_field(g, 0); }
void orc_ade_h(orc_grid *g) { ade_field(g, 1); }

/* --------------------------------------------------------------------------
 * First-order Mur, AdvanceMagneticMUR (bordersmur.F90:1129-1383, the
 * "first order" branch; the second-order one stops with an error in the
 * reference, :519).  On each absorbing face the two tangential H components
 * of the plane just outside their sweep are advanced from the first plane
 * inside:  H_b = Past_1 + CAB1[m] (H_1 - Past_b), m = media index of the
 * inside cell; then (:1388-1600) Past <- H on both planes.  The face regions
 * span the components' sweeps along the two other axes (:113-160); the
 * updated cells of different faces never coincide and never feed each other,
 * so one Past array per component serves every face.
 * ------------------------------------------------------------------------ */
void orc_mur_table(int32_t nmedia, const real *epr, const real *mur, double dt, real eps0,
                   real mu0, real idh, real *cab1) {
  const real cluz = R(1.0) / RSQRT(eps0 * mu0);
  for (int32_t m = 0; m <= nmedia; ++m) {
    const real v = RSQRT(epr[m] * mur[m]);
    const real cnum = (real)((double)(R(1.0) / idh) / (dt * (double)cluz / (double)v));
    cab1[m] = (R(1.0) - cnum) / (R(1.0) + cnum);
  }
}

/* the tangential H components of face f (0 back .. 5 up) in the reference's
 * loop order, and the face's normal axis and side */
static const int mur_comp[6][2] = {{ORC_HZ, ORC_HY}, {ORC_HZ, ORC_HY}, {ORC_HX, ORC_HZ},
                                   {ORC_HX, ORC_HZ}, {ORC_HY, ORC_HX}, {ORC_HY, ORC_HX}};

static void mur_plane(orc_grid *g, int f, int c, int *ghost, int *inner) {
  const int a = f / 2, hi = f & 1;
  const int last = sweep_hi(g, c, a);
  *ghost = hi ? last + 1 : -1;
  *inner = hi ? last : 0;
}

void orc_mur_h(orc_grid *g) {
  /* reference face order: left, right, down, up, back, front */
  static const int order[6] = {2, 3, 4, 5, 0, 1};
  for (int o = 0; o < 6; ++o) {
    const int f = order[o];
    if (!g->mur[f]) continue;
    const int a = f / 2;
    for (int t = 0; t < 2; ++t) {
      const int c = mur_comp[f][t];
      int ghost, inner;
      mur_plane(g, f, c, &ghost, &inner);
      real *H = g->f[c];
      real *P = g->murpast[c - 3];
      const int hi0 = sweep_hi(g, c, 0), hi1 = sweep_hi(g, c, 1), hi2 = sweep_hi(g, c, 2);
      for (int k = 0; k <= hi2; ++k)
        for (int j = 0; j <= hi1; ++j)
          for (int i = 0; i <= hi0; ++i) {
            int cb[3] = {i, j, k}, ci[3] = {i, j, k};
            if ((a == 0 && i != 0) || (a == 1 && j != 0) || (a == 2 && k != 0)) continue;
            cb[a] = ghost;
            ci[a] = inner;
            const size_t pb = IDX(g, cb[0], cb[1], cb[2]);
            const size_t pi = IDX(g, ci[0], ci[1], ci[2]);
            const int m = g->mi[c][pi];
            H[pb] = P[pi] + g->cab1[f][m] * (H[pi] - P[pb]);
          }
    }
  }
  for (int o = 0; o < 6; ++o) {
    const int f = order[o];
    if (!g->mur[f]) continue;
    const int a = f / 2;
    for (int t = 0; t < 2; ++t) {
      const int c = mur_comp[f][t];
      int ghost, inner;
      mur_plane(g, f, c, &ghost, &inner);
      real *H = g->f[c];
      real *P = g->murpast[c - 3];
      const int hi0 = sweep_hi(g, c, 0), hi1 = sweep_hi(g, c, 1), hi2 = sweep_hi(g, c, 2);
      for (int k = 0; k <= hi2; ++k)
        for (int j = 0; j <= hi1; ++j)
          for (int i = 0; i <= hi0; ++i) {
            if ((a == 0 && i != 0) || (a == 1 && j != 0) || (a == 2 && k != 0)) continue;
            for (int w = 0; w < 2; ++w) {
              int cc[3] = {i, j, k};
              cc[a] = w ? inner : ghost;
              const size_t p = IDX(g, cc[0], cc[1], cc[2]);
              P[p] = H[p];
            }
          }
    }
  }
}

/* --------------------------------------------------------------------------
 * Nodal sources: evolucion (linear interpolation of the excitation table,
 * nodalsources.F90:317-356) and AdvancenodalE/H (:396-578, :584-777).
 * E time stamp tiempo(n) cast to the real kind (:426); H stamp
 * tiempo(n)+0.5*dt (:618).  Hard: F = amp*f(t); soft: F = F - G2*Id_a*Id_b*
 * amp*f(t) with the reference's factor order.  PEC (E) / PMC (H) media are
 * skipped; the H hard sources read the Hx media matrix for Hy and Hz as the
 * reference does (:641,:715).
 * ------------------------------------------------------------------------ */
real orc_evolucion(real t, const orc_nodal *s) {
  real de = s->deltaevol;
  long long nprev = (long long)(t / de);
  if (nprev + 1 > s->numus || nprev + 1 <= 0) return R(0.0);
  return (s->evol[nprev + 1] - s->evol[nprev]) / de * (t - (real)nprev * de) +
         s->evol[nprev];
}

static void nodal_apply(orc_grid *g, const orc_nodal *s, real t) {
  const int c = s->comp;
  real *F = g->f[c];
  const int16_t *mi = g->mi[(c >= 4 && s->hard) ? 3 : c];
  const int skipbit = c < 3 ? ORC_KIND_PEC : ORC_KIND_PMC;
  const real *gg = c < 3 ? g->g2 : g->gm2;
  const real *Ix = c < 3 ? g->id[3] : g->id[0];
  const real *Iy = c < 3 ? g->id[4] : g->id[1];
  const real *Iz = c < 3 ? g->id[5] : g->id[2];
  const int cc = c % 3;
  for (int k = s->zi; k <= s->ze; ++k)
    for (int j = s->yi; j <= s->ye; ++j)
      for (int i = s->xi; i <= s->xe; ++i) {
        size_t p = IDX(g, i, j, k);
        int m = mi[p];
        if (g->medkind[m] & skipbit) continue;
        if (s->hard) {
          F[p] = s->amp * orc_evolucion(t, s);
        } else {
          real ia, ib;
          if (cc == 0) { ia = Iy[j + 1]; ib = Iz[k + 1]; }
          else if (cc == 1) { ia = Ix[i + 1]; ib = Iz[k + 1]; }
          else { ia = Iy[j + 1]; ib = Ix[i + 1]; }
          F[p] = F[p] - gg[m] * ia * ib * s->amp * orc_evolucion(t, s);
        }
      }
}

void orc_nodal_e(orc_grid *g, int32_t n) {
  real t = (real)g->tiempo[n];
  for (int c = 0; c < 3; ++c)
    for (int hard = 1; hard >= 0; --hard)
      for (int s = 0; s < g->nnodal; ++s)
        if (g->nodal[s].comp == c && g->nodal[s].hard == hard)
          nodal_apply(g, &g->nodal[s], t);
}

void orc_nodal_h(orc_grid *g, int32_t n) {
  real t = (real)(g->tiempo[n] + (double)R(0.5) * g->dt);
  for (int c = 3; c < 6; ++c)
    for (int hard = 1; hard >= 0; --hard)
      for (int s = 0; s < g->nnodal; ++s)
        if (g->nodal[s].comp == c && g->nodal[s].hard == hard)
          nodal_apply(g, &g->nodal[s], t);
}

/* --------------------------------------------------------------------------
 * TF/SF plane waves, planewaves.F90.
 * orc_lines restates the discretisation lines of preprocess_geom.F90:
 * 5330-5347 (interior: line(i+1) = d(i) + line(i), origin 0) and 5690-5712
 * (CPML and padding extrapolated with the first/last interior cell).
 * ------------------------------------------------------------------------ */
void orc_lines(int32_t n, int32_t Llo, int32_t Lhi, const real *DX, real *line) {
  const int lo = Llo, hi = n - Lhi;            /* interior nodes lo..hi */
  line[lo + 1] = R(0.0);
  for (int i = lo; i < hi; ++i) line[i + 2] = DX[i + 2] + line[i + 1];
  real delta = line[lo + 2] - line[lo + 1];
  for (int i = lo - 1; i >= -1; --i) line[i + 1] = line[i + 2] - delta;
  delta = line[hi + 1] - line[hi];
  for (int i = hi + 1; i <= n + 1; ++i) line[i + 1] = line[i] + delta;
}

/* face limits of one component (ehxyz entries of TrFr/IzDe/AbAr) */
typedef struct { int lo[3], hi[3]; int fixed_lo, fixed_hi; } pw_face;

struct orc_pw_state {
  int npw;
  real cluz;
  real *phys[6][3];               /* PhysCoor(field)%x/y/z over [-1, n+1] */
  struct {
    real fpw[6][ORC_PW_MODES];    /* E and H amplitudes per mode */
    real pn[ORC_PW_MODES][3];     /* normalised direction */
    real dist0[ORC_PW_MODES];     /* distanciaInicial */
    int ilum[6];                  /* Tr Fr Iz De Ab Ar */
    pw_face tr[6], iz[6], ab[6];  /* by component: x faces, y faces, z faces */
  } *w;
};

static int imax(int a, int b) { return a > b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }

/* SINPMLSweep of component c along axis a (the sweep without the CPML) */
static int sin_lo(const orc_grid *g, int c, int a) { (void)c; return g->L[a][0]; }
static int sin_hi(const orc_grid *g, int c, int a) { return sweep_hi(g, c, a) - g->L[a][1]; }

/* InitPlaneWave, planewaves.F90:72-766 */
int orc_pw_init(orc_grid *g) {
  orc_pw_free(g);
  if (g->npw <= 0) return 0;
  orc_pw_state *S = (orc_pw_state *)calloc(1, sizeof(orc_pw_state));
  if (!S) return -1;
  S->npw = g->npw;
  S->w = calloc((size_t)g->npw, sizeof(*S->w));
  const int n3[3] = {g->nx, g->ny, g->nz};
  /* cluz and zvac in the real kind (:91-92) */
  S->cluz = R(1.0) / RSQRT(g->eps0 * g->mu0);
  const real zvac = RSQRT(g->mu0 / g->eps0);
  /* physical coordinates of every component (:94-164): node axes read the
   * line, cell axes the midpoint of two lines */
  for (int c = 0; c < 6; ++c)
    for (int a = 0; a < 3; ++a) {
      real *v = (real *)calloc((size_t)n3[a] + 3, sizeof(real));
      S->phys[c][a] = v;
      const int node = c < 3 ? (c != a) : (c - 3 == a);
      const int top = sweep_hi(g, c, a) + (node ? 1 : 0);
      for (int i = -1; i <= top; ++i)
        v[i + 1] = node ? g->line[a][i + 1]
                        : (g->line[a][i + 1] + g->line[a][i + 2]) * R(0.5);
    }
  for (int q = 0; q < g->npw; ++q) {
    const orc_planewave *P = &g->pw[q];
    const int *E = P->esq;  /* x1 x2 y1 y2 z1 z2 */
    /* illuminated faces (:203-214), against the H sweeps */
    S->w[q].ilum[0] = E[0] >= sin_lo(g, ORC_HX, 0) && E[0] <= sin_hi(g, ORC_HX, 0);
    S->w[q].ilum[1] = E[1] <= sin_hi(g, ORC_HX, 0) && E[1] >= sin_lo(g, ORC_HX, 0);
    S->w[q].ilum[2] = E[2] >= sin_lo(g, ORC_HY, 1) && E[2] <= sin_hi(g, ORC_HY, 1);
    S->w[q].ilum[3] = E[3] <= sin_hi(g, ORC_HY, 1) && E[3] >= sin_lo(g, ORC_HY, 1);
    S->w[q].ilum[4] = E[4] >= sin_lo(g, ORC_HZ, 2) && E[4] <= sin_hi(g, ORC_HZ, 2);
    S->w[q].ilum[5] = E[5] <= sin_hi(g, ORC_HZ, 2) && E[5] >= sin_lo(g, ORC_HZ, 2);
    /* face limits (:216-295).  For the x faces of Ez/Ey, the y/z range of a
     * component stops one short of the box end along its own axis. */
    pw_face *tr = S->w[q].tr, *iz = S->w[q].iz, *ab = S->w[q].ab;
    for (int c = 0; c < 3; ++c) {
      for (int a = 0; a < 3; ++a) {
        const int lo = imax(sin_lo(g, c, a), E[2 * a]);
        const int hi_box = E[2 * a + 1] - (a == c ? 1 : 0);
        const int hi = imin(sin_hi(g, c, a), hi_box);
        tr[c].lo[a] = iz[c].lo[a] = ab[c].lo[a] = lo;
        tr[c].hi[a] = iz[c].hi[a] = ab[c].hi[a] = hi;
      }
      tr[c].fixed_lo = imax(sin_lo(g, c, 0), E[0]);
      tr[c].fixed_hi = imin(sin_hi(g, c, 0), E[1]);
      iz[c].fixed_lo = imax(sin_lo(g, c, 1), E[2]);
      iz[c].fixed_hi = imin(sin_hi(g, c, 1), E[3]);
      ab[c].fixed_lo = imax(sin_lo(g, c, 2), E[4]);
      ab[c].fixed_hi = imin(sin_hi(g, c, 2), E[5]);
    }
    /* H faces copy the E face they pair with, the low face one cell lower:
     * Hy<-Ez and Hz<-Ey on x faces, Hz<-Ex and Hx<-Ez on y faces, Hx<-Ey and
     * Hy<-Ex on z faces */
    tr[ORC_HY] = tr[ORC_EZ]; tr[ORC_HY].fixed_lo -= 1;
    tr[ORC_HZ] = tr[ORC_EY]; tr[ORC_HZ].fixed_lo -= 1;
    iz[ORC_HZ] = iz[ORC_EX]; iz[ORC_HZ].fixed_lo -= 1;
    iz[ORC_HX] = iz[ORC_EZ]; iz[ORC_HX].fixed_lo -= 1;
    ab[ORC_HX] = ab[ORC_EY]; ab[ORC_HX].fixed_lo -= 1;
    ab[ORC_HY] = ab[ORC_EX]; ab[ORC_HY].fixed_lo -= 1;
    /* directions, amplitudes (:330-342) and H amplitudes (:1494-1509) */
    for (int m = 0; m < P->nmodes; ++m) {
      const real px = P->p[m][0], py = P->p[m][1], pz = P->p[m][2];
      const real modulus = RSQRT(px * px + py * py + RPOW(pz, R(2.0)));
      real *pn = S->w[q].pn[m];
      pn[0] = px / modulus; pn[1] = py / modulus; pn[2] = pz / modulus;
      S->w[q].fpw[0][m] = P->e[m][0];
      S->w[q].fpw[1][m] = P->e[m][1];
      S->w[q].fpw[2][m] = P->e[m][2];
      S->w[q].fpw[3][m] = (pn[1] * P->e[m][2] - pn[2] * P->e[m][1]) / zvac;
      S->w[q].fpw[4][m] = (pn[2] * P->e[m][0] - pn[0] * P->e[m][2]) / zvac;
      S->w[q].fpw[5][m] = (pn[0] * P->e[m][1] - pn[1] * P->e[m][0]) / zvac;
      /* null-phase corner (:372-407): per axis the low corner when the
       * direction is >= 0, else the high one, clipped to the non-CPML region */
      real c0[3];
      const int hc[3] = {ORC_HX, ORC_HY, ORC_HZ};
      real span2[3];
      for (int a = 0; a < 3; ++a) {
        const int clo = imax(E[2 * a] - 1, sin_lo(g, hc[a], a));
        const int chi = imin(E[2 * a + 1] + 1, sin_hi(g, hc[a], a));
        c0[a] = pn[a] >= R(0.0) ? g->line[a][clo + 1] : g->line[a][chi + 1];
        const real dd = g->line[a][clo + 1] - g->line[a][chi + 1];
        span2[a] = RPOW(dd, R(2.0));
      }
      const real diag = RSQRT(span2[0] + span2[1] + span2[2]);
      S->w[q].dist0[m] = (c0[0] * pn[0] + c0[1] * pn[1] + c0[2] * pn[2]) - P->incert[m] * diag;
    }
  }
  g->pws = S;
  return 0;
}

void orc_pw_free(orc_grid *g) {
  orc_pw_state *S = g->pws;
  if (!S) return;
  for (int c = 0; c < 6; ++c)
    for (int a = 0; a < 3; ++a) free(S->phys[c][a]);
  free(S->w);
  free(S);
  g->pws = NULL;
}

/* evolucion of the plane wave (:824-851): zero before the second sample and
 * after the table */
static real pw_evol(const orc_planewave *P, real cluz, real t, real d) {
  real ev = R(0.0);
  long long nprev = (long long)(int)((t - d / cluz) / P->deltaevol);
  if (nprev + 1 <= P->numus) {
    if (nprev > 0)
      ev = (P->evol[nprev + 1] - P->evol[nprev]) / P->deltaevol *
               ((t - d / cluz) - (real)nprev * P->deltaevol) + P->evol[nprev];
  }
  return ev;
}

/* Incid (:778-853) for plane wave q, component nf at (i,j,k) */
static real pw_incid(const orc_grid *g, int q, int nf, real t, int i, int j, int k) {
  const orc_pw_state *S = g->pws;
  const orc_planewave *P = &g->pw[q];
  const real xf = S->phys[nf][0][i + 1], yf = S->phys[nf][1][j + 1], zf = S->phys[nf][2][k + 1];
  real ehi = R(0.0);
  for (int m = 0; m < P->nmodes; ++m) {
    const real *pn = S->w[q].pn[m];
    const real d = (xf * pn[0] + yf * pn[1] + zf * pn[2]) - S->w[q].dist0[m];
    ehi = ehi + S->w[q].fpw[nf][m] * pw_evol(P, S->cluz, t, d);
  }
  return ehi;
}

/* one face of one component: F(pos) += sign * G * Incid(nf, pos + off) * Id(fixed) */
static void pw_face_apply(orc_grid *g, int q, int comp, int axis, int fixed,
                          const pw_face *fc, int nf, int off, real sign, real G,
                          const real *Id, real t) {
  real *F = g->f[comp];
  const real id = Id[fixed + 1];
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) { lo[a] = fc->lo[a]; hi[a] = fc->hi[a]; }
  lo[axis] = hi[axis] = fixed;
  for (int k = lo[2]; k <= hi[2]; ++k)
    for (int j = lo[1]; j <= hi[1]; ++j)
      for (int i = lo[0]; i <= hi[0]; ++i) {
        const int ii = i + (axis == 0 ? off : 0);
        const int jj = j + (axis == 1 ? off : 0);
        const int kk = k + (axis == 2 ? off : 0);
        const real inc = pw_incid(g, q, nf, t, ii, jj, kk);
        const size_t p = IDX(g, i, j, k);
        if (sign > R(0.0)) F[p] = F[p] + G * inc * id;
        else F[p] = F[p] - G * inc * id;
      }
}

/* AdvancePlaneWaveE, planewaves.F90:878-1174 (G2 of the background medium 1) */
void orc_pw_e(orc_grid *g, int32_t n) {
  if (!g->pws) return;
  const real t = (real)g->tiempo[n];
  const real G = g->g2[1];
  const real *Ix = g->id[3], *Iy = g->id[4], *Iz = g->id[5];
  for (int q = 0; q < g->npw; ++q) {
    const int *il = g->pws->w[q].ilum;
    const pw_face *tr = g->pws->w[q].tr, *iz = g->pws->w[q].iz, *ab = g->pws->w[q].ab;
    if (il[0]) {  /* back */
      pw_face_apply(g, q, ORC_EZ, 0, tr[ORC_EZ].fixed_lo, &tr[ORC_EZ], ORC_HY, -1, R(-1), G, Ix, t);
      pw_face_apply(g, q, ORC_EY, 0, tr[ORC_EY].fixed_lo, &tr[ORC_EY], ORC_HZ, -1, R(1), G, Ix, t);
    }
    if (il[1]) {  /* front */
      pw_face_apply(g, q, ORC_EZ, 0, tr[ORC_EZ].fixed_hi, &tr[ORC_EZ], ORC_HY, 0, R(1), G, Ix, t);
      pw_face_apply(g, q, ORC_EY, 0, tr[ORC_EY].fixed_hi, &tr[ORC_EY], ORC_HZ, 0, R(-1), G, Ix, t);
    }
    if (il[2]) {  /* left */
      pw_face_apply(g, q, ORC_EX, 1, iz[ORC_EX].fixed_lo, &iz[ORC_EX], ORC_HZ, -1, R(-1), G, Iy, t);
      pw_face_apply(g, q, ORC_EZ, 1, iz[ORC_EZ].fixed_lo, &iz[ORC_EZ], ORC_HX, -1, R(1), G, Iy, t);
    }
    if (il[3]) {  /* right */
      pw_face_apply(g, q, ORC_EZ, 1, iz[ORC_EZ].fixed_hi, &iz[ORC_EZ], ORC_HX, 0, R(-1), G, Iy, t);
      pw_face_apply(g, q, ORC_EX, 1, iz[ORC_EX].fixed_hi, &iz[ORC_EX], ORC_HZ, 0, R(1), G, Iy, t);
    }
    if (il[4]) {  /* down */
      pw_face_apply(g, q, ORC_EX, 2, ab[ORC_EX].fixed_lo, &ab[ORC_EX], ORC_HY, -1, R(1), G, Iz, t);
      pw_face_apply(g, q, ORC_EY, 2, ab[ORC_EY].fixed_lo, &ab[ORC_EY], ORC_HX, -1, R(-1), G, Iz, t);
    }
    if (il[5]) {  /* up */
      pw_face_apply(g, q, ORC_EX, 2, ab[ORC_EX].fixed_hi, &ab[ORC_EX], ORC_HY, 0, R(-1), G, Iz, t);
      pw_face_apply(g, q, ORC_EY, 2, ab[ORC_EY].fixed_hi, &ab[ORC_EY], ORC_HX, 0, R(1), G, Iz, t);
    }
  }
}

/* AdvancePlaneWaveH, planewaves.F90:1180-1475 (GM2 of medium 1), at
 * tiempo(n) + dt/2 */
void orc_pw_h(orc_grid *g, int32_t n) {
  if (!g->pws) return;
  const real t = (real)(g->tiempo[n] + (double)R(0.5) * g->dt);
  const real G = g->gm2[1];
  const real *Ix = g->id[0], *Iy = g->id[1], *Iz = g->id[2];
  for (int q = 0; q < g->npw; ++q) {
    const int *il = g->pws->w[q].ilum;
    const pw_face *tr = g->pws->w[q].tr, *iz = g->pws->w[q].iz, *ab = g->pws->w[q].ab;
    if (il[0]) {
      pw_face_apply(g, q, ORC_HZ, 0, tr[ORC_HZ].fixed_lo, &tr[ORC_HZ], ORC_EY, 1, R(1), G, Ix, t);
      pw_face_apply(g, q, ORC_HY, 0, tr[ORC_HY].fixed_lo, &tr[ORC_HY], ORC_EZ, 1, R(-1), G, Ix, t);
    }
    if (il[1]) {
      pw_face_apply(g, q, ORC_HZ, 0, tr[ORC_HZ].fixed_hi, &tr[ORC_HZ], ORC_EY, 0, R(-1), G, Ix, t);
      pw_face_apply(g, q, ORC_HY, 0, tr[ORC_HY].fixed_hi, &tr[ORC_HY], ORC_EZ, 0, R(1), G, Ix, t);
    }
    if (il[2]) {
      pw_face_apply(g, q, ORC_HX, 1, iz[ORC_HX].fixed_lo, &iz[ORC_HX], ORC_EZ, 1, R(1), G, Iy, t);
      pw_face_apply(g, q, ORC_HZ, 1, iz[ORC_HZ].fixed_lo, &iz[ORC_HZ], ORC_EX, 1, R(-1), G, Iy, t);
    }
    if (il[3]) {
      pw_face_apply(g, q, ORC_HX, 1, iz[ORC_HX].fixed_hi, &iz[ORC_HX], ORC_EZ, 0, R(-1), G, Iy, t);
      pw_face_apply(g, q, ORC_HZ, 1, iz[ORC_HZ].fixed_hi, &iz[ORC_HZ], ORC_EX, 0, R(1), G, Iy, t);
    }
    if (il[4]) {
      pw_face_apply(g, q, ORC_HX, 2, ab[ORC_HX].fixed_lo, &ab[ORC_HX], ORC_EY, 1, R(-1), G, Iz, t);
      pw_face_apply(g, q, ORC_HY, 2, ab[ORC_HY].fixed_lo, &ab[ORC_HY], ORC_EX, 1, R(1), G, Iz, t);
    }
    if (il[5]) {
      pw_face_apply(g, q, ORC_HX, 2, ab[ORC_HX].fixed_hi, &ab[ORC_HX], ORC_EY, 0, R(1), G, Iz, t);
      pw_face_apply(g, q, ORC_HY, 2, ab[ORC_HY].fixed_hi, &ab[ORC_HY], ORC_EX, 0, R(-1), G, Iz, t);
    }
  }
}

/* --------------------------------------------------------------------------
 * Frequency-domain probes, observation.F90:3527-3600.  mcPI2 is formed in
 * real(4) and stored in complex(8) (fdetypes.F90:187-190): (-0, -2*pi_4).
 * exp of a complex with a (signed) zero real part is (cos y, sin y); the
 * products of a complex(8) by a real factor reduce to scaling both parts.
 * ------------------------------------------------------------------------ */
void orc_dft_aux(int32_t nfreq, const float *freq, double dt, double at, int32_t half,
                 double *aux) {
  const float pi4 = 3.14159265358979323846f;
  const double mc = -(double)(2.0f * pi4);          /* imaginary part of mcPI2 */
  for (int f = 0; f < nfreq; ++f) {
    const double F = (double)freq[f];
    const double y = (mc * F) * at;
    double re = dt * cos(y), im = dt * sin(y);       /* (dt, 0) * exp(j y) */
    if (half) {
      const double y2 = ((mc * F) * dt) * 0.5;
      const double c = cos(y2), s = sin(y2);
      const double r2 = re * c - im * s, i2 = re * s + im * c;
      re = r2;
      im = i2;
    }
    aux[2 * f] = re;
    aux[2 * f + 1] = im;
  }
}

static int dft_count(int lo, int hi, int st) {
  int c = 0;
  for (int v = lo; v <= hi; ++v) c += (v % st == 0);
  return c;
}

void orc_dft_accum(const orc_grid *g, int32_t kind, const int32_t box[6],
                   const int32_t stride[3], int32_t nfreq, const double *aux, double *acc) {
  const int ni = dft_count(box[0], box[1], stride[0]);
  const int nj = dft_count(box[2], box[3], stride[1]);
  int kt = 0;
  for (int k = box[4]; k <= box[5]; ++k) {
    if (k % stride[2]) continue;
    int jt = 0;
    for (int j = box[2]; j <= box[3]; ++j) {
      if (j % stride[1]) continue;
      int it = 0;
      for (int i = box[0]; i <= box[1]; ++i) {
        if (i % stride[0]) continue;
        const size_t pt = ((size_t)kt * nj + jt) * ni + it;
        for (int c = 0; c < 3; ++c) {
          const double v = (double)g->f[3 * kind + c][IDX(g, i, j, k)];
          double *a = acc + (pt * 3 + c) * (size_t)nfreq * 2;
          for (int f = 0; f < nfreq; ++f) {
            a[2 * f] = a[2 * f] + aux[2 * f] * v;
            a[2 * f + 1] = a[2 * f + 1] + aux[2 * f + 1] * v;
          }
        }
        ++it;
      }
      ++jt;
    }
    ++kt;
  }
}

/* One full step in the order of ciclo_temporal, timestepping.F90:1248-1981:
 * E Yee -> E CPML -> E ADE -> plane wave E -> nodal E -> H Yee -> H CPML ->
 * H ADE -> plane wave H -> nodal H -> Mur (:1286-1441, :1497-1615,
 * :1585-1586, :1711-1712). */
void orc_step(orc_grid *g, int32_t n) {
  orc_yee_e(g);
  orc_cpml_e(g);
  if (g->ndisp) orc_ade_e(g);
  if (g->pws) orc_pw_e(g, n);
  if (g->nnodal) orc_nodal_e(g, n);
  orc_yee_h(g);
  orc_cpml_h(g);
  if (g->nmdisp) orc_ade_h(g);
  if (g->pws) orc_pw_h(g, n);
  if (g->nnodal) orc_nodal_h(g, n);
  if (g->mur[0] || g->mur[1] || g->mur[2] || g->mur[3] || g->mur[4] || g->mur[5])
    orc_mur_h(g);
}

void orc_run(orc_grid *g, int32_t n0, int32_t nsteps, int32_t nprobe, const int32_t *pc,
             const int32_t *pi, const int32_t *pj, const int32_t *pk, real *out) {
  for (int32_t s = 0; s < nsteps; ++s) {
    orc_step(g, n0 + s);
    for (int32_t p = 0; p < nprobe; ++p)
      out[(size_t)s * nprobe + p] = g->f[pc[p]][IDX(g, pi[p], pj[p], pk[p])];
  }
}
