/*
 * fdtd_oracle.h -- CPU restatement of semba-fdtd's field time-stepping path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the HIP library is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product path never does.
 *
 * Layout shared with the HIP library (and with the reference's Alloc box,
 * semba_fdtd.F90:1285-1300): every field / media-index array covers the cell
 * coordinates [-1, n+1] on each axis, x fastest:
 *     idx(i,j,k) = ((k+1)*(ny+3) + (j+1))*(nx+3) + (i+1)
 * Coordinates are "total" cells: 0 is the outer face of the low CPML layer,
 * nx/ny/nz count the interior plus both CPML slabs.  1-D spacing tables cover
 * [-1, n+1] (index i+1); the raw cell-size array DX covers [-2, n+1]
 * (index i+2), as sgg%dx does (preprocess_geom.F90:5639-5666).
 *
 * orc_real is the reference's RKIND: real(4) by default, real(8) when built
 * with -DORC_REAL64 (the FP64 yardstick).
 */
#ifndef FDTD_ORACLE_H
#define FDTD_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifdef ORC_REAL64
typedef double orc_real;
#else
typedef float orc_real;
#endif

#define ORC_PMAX 8 /* max poles per dispersive medium */

enum { ORC_EX = 0, ORC_EY, ORC_EZ, ORC_HX, ORC_HY, ORC_HZ };
/* psi slots: E then H; (component, stretched axis) */
enum {
  ORC_PSI_EXY = 0, ORC_PSI_EXZ, ORC_PSI_EYX, ORC_PSI_EYZ, ORC_PSI_EZX, ORC_PSI_EZY,
  ORC_PSI_HXY, ORC_PSI_HXZ, ORC_PSI_HYX, ORC_PSI_HYZ, ORC_PSI_HZX, ORC_PSI_HZY
};

typedef struct {
  int32_t comp;        /* ORC_EX..ORC_HZ */
  int32_t hard;        /* 1: E = amp*f(t) ; 0: soft current E -= G2*Id*Id*amp*f(t) */
  int32_t xi, xe, yi, ye, zi, ze; /* box, already clipped to the sweep */
  orc_real amp;
  orc_real deltaevol;  /* sample spacing of the excitation table */
  int32_t numus;       /* samples are evol[0..numus] */
  const orc_real *evol;
} orc_nodal;

#define ORC_PW_MODES 4 /* max modes per plane wave */

/* One TF/SF plane-wave box (planeonde_t, fdetypes.F90:458-465) */
typedef struct {
  int32_t esq[6];      /* esqx1, esqx2, esqy1, esqy2, esqz1, esqz2 (total cells) */
  int32_t nmodes;
  orc_real p[ORC_PW_MODES][3];   /* propagation direction (normalised at init) */
  orc_real e[ORC_PW_MODES][3];   /* E amplitude */
  orc_real incert[ORC_PW_MODES];
  orc_real deltaevol;
  int32_t numus;
  const orc_real *evol;          /* evol[0..numus] */
} orc_planewave;

typedef struct orc_pw_state orc_pw_state;

typedef struct {
  int32_t nx, ny, nz;
  orc_real *f[6];              /* Ex Ey Ez Hx Hy Hz, Alloc layout */
  const int16_t *mi[6];        /* media index per component, Alloc layout */
  const orc_real *g1, *g2, *gm1, *gm2; /* 0..nmedia */
  int32_t nmedia;
  const orc_real *id[6];       /* Idxe Idye Idze Idxh Idyh Idzh, length n+3 */
  /* CPML */
  int32_t L[3][2];             /* layers (axis, side) ; 0 = none */
  const orc_real *pbe[3], *pce[3], *pbm[3], *pcm[3]; /* per axis, length n+3 */
  orc_real *psi[12];           /* full Alloc-sized arrays, only slab entries used */
  /* electric pole-residue ADE */
  int32_t ndisp;
  const int32_t *disp_medium;  /* [ndisp] media index */
  const int32_t *disp_npoles;  /* [ndisp] */
  const double *disp_kappa;    /* [ndisp][ORC_PMAX][2] */
  const double *disp_beta;     /* [ndisp][ORC_PMAX][2] */
  const double *disp_g3;       /* [ndisp][ORC_PMAX][2] */
  orc_real *eprev[3];          /* Alloc-sized, per E component */
  double *jcur[3];             /* Alloc-sized * jstride * 2 per E component */
  int32_t jstride;             /* pole slots per edge (>= max poles) */
  double dt;
  const double *tiempo;        /* tiempo[n], n = 0..nsteps+1 */
  int32_t nnodal;
  const orc_nodal *nodal;
  const int32_t *medkind;      /* [nmedia+1] bits: 1 PEC, 2 PMC, 4 E-dispersive */
  /* TF/SF plane waves */
  int32_t npw;
  const orc_planewave *pw;
  const orc_real *line[3];     /* node coordinates LineX/Y/Z over [-1, n+1] */
  orc_real eps0, mu0;
  orc_pw_state *pws;           /* derived by orc_pw_init */
  /* magnetic pole-residue ADE (magneticdispersive.F90), the E layout on H */
  int32_t nmdisp;
  const int32_t *mdisp_medium; /* [nmdisp] media index */
  const int32_t *mdisp_npoles; /* [nmdisp] */
  const double *mdisp_kappa;   /* [nmdisp][ORC_PMAX][2] */
  const double *mdisp_beta;    /* [nmdisp][ORC_PMAX][2] */
  const double *mdisp_gm3;     /* [nmdisp][ORC_PMAX][2] */
  orc_real *hprev[3];          /* Alloc-sized, per H component */
  double *mjcur[3];            /* Alloc-sized * jstride * 2 per H component */
  /* first-order Mur faces (bordersmur.F90), order back, front, left, right,
   * down, up (x-, x+, y-, y+, z-, z+) */
  int32_t mur[6];              /* face absorbs */
  const orc_real *cab1[6];     /* [nmedia+1] per face */
  orc_real *murpast[3];        /* Alloc-sized per H component: the field of the
                                  previous step on the Mur planes */
} orc_grid;

/* ---- coefficient tables ---- */
double orc_dt(orc_real dxmin, orc_real dymin, orc_real dzmin, orc_real eps0, orc_real mu0,
              orc_real cfl, int32_t forcecfl, double dt_in);
void orc_g_tables(int32_t nmedia, const orc_real *epr, const orc_real *sigma,
                  const orc_real *mur, const orc_real *sigmam, const int32_t *kind,
                  double dt, orc_real eps0, orc_real mu0,
                  orc_real *g1, orc_real *g2, orc_real *gm1, orc_real *gm2);
void orc_spacing(int32_t n, const orc_real *DX, orc_real *de, orc_real *dh,
                 orc_real *Ide, orc_real *Idh);
void orc_cpml_axis(int32_t n, int32_t Llo, int32_t Lhi,
                   orc_real orden_lo, orc_real orden_hi, orc_real R_lo, orc_real R_hi,
                   orc_real alphamaxpar, orc_real alphaorden, orc_real kappamaxpar,
                   double dt, orc_real eps0, orc_real mu0,
                   const orc_real *de, const orc_real *dh,
                   orc_real *pbe, orc_real *pce, orc_real *pbm, orc_real *pcm,
                   orc_real *Ide, orc_real *Idh);
void orc_ade_tables(int32_t npoles, const double *c_re_im, const double *a_re_im,
                    orc_real eps11, orc_real sigma11, double dt,
                    double *kappa, double *beta, double *g3,
                    orc_real *g1_out, orc_real *g2_out);
void orc_time_vector(int32_t nlast, double dt, double *tiempo);
/* First-order Mur coefficient CAB1 of one face for every medium
 * (calc_murconstants, bordersmur.F90:316-363): idh = the dual inverse
 * spacing at the face's end of the Alloc range (Idxh(XI) / Idxh(XE) ...);
 * cnum = (1/idh) / (dt * c0/sqrt(epr*mur)), CAB1 = (1-cnum)/(1+cnum). */
void orc_mur_table(int32_t nmedia, const orc_real *epr, const orc_real *mur, double dt,
                   orc_real eps0, orc_real mu0, orc_real idh, orc_real *cab1);

/* ---- stepping ---- */
void orc_yee_e(orc_grid *g);
void orc_yee_h(orc_grid *g);
void orc_cpml_e(orc_grid *g);
void orc_cpml_h(orc_grid *g);
void orc_ade_e(orc_grid *g);
void orc_ade_h(orc_grid *g);
void orc_mur_h(orc_grid *g);
orc_real orc_evolucion(orc_real t, const orc_nodal *s);
void orc_nodal_e(orc_grid *g, int32_t n);
void orc_nodal_h(orc_grid *g, int32_t n);
/* TF/SF plane waves (planewaves.F90): orc_pw_init derives the constants,
 * coordinates and face limits of InitPlaneWave; orc_pw_e / orc_pw_h are
 * AdvancePlaneWaveE / AdvancePlaneWaveH at step n. */
int orc_pw_init(orc_grid *g);
void orc_pw_free(orc_grid *g);
void orc_pw_e(orc_grid *g, int32_t n);
void orc_pw_h(orc_grid *g, int32_t n);
/* node coordinates of one axis (preprocess_geom.F90:5330-5347, 5690-5712):
 * origin 0 at the first interior node (total index Llo), cumulative interior
 * cell sizes, CPML and padding extrapolated with the edge interior cells;
 * DX over [-2, n+1], line over [-1, n+1]. */
void orc_lines(int32_t n, int32_t Llo, int32_t Lhi, const orc_real *DX, orc_real *line);
/* Frequency-domain observation (observation.F90:3527-3600, 3532-3533):
 * orc_dft_aux fills aux[f] = dt * exp(mcPI2 * freq[f] * at) for E (half 0,
 * at = tiempo(n)) or that times exp(mcPI2 * freq[f] * dt * 0.5) for H (half 1),
 * with mcPI2 = -j*2*pi formed in real(4) (fdetypes.F90:190); aux is
 * [nfreq][re,im].  orc_dft_accum adds aux * F for the three components of
 * kind (0: Ex,Ey,Ez; 1: Hx,Hy,Hz) at the points i1..i2 step si (and j, k) whose
 * coordinates are multiples of the stride, into acc
 * [k'][j'][i'][comp][freq][re,im] (valor3DComplex's Fortran order). */
void orc_dft_aux(int32_t nfreq, const float *freq, double dt, double at, int32_t half,
                 double *aux);
void orc_dft_accum(const orc_grid *g, int32_t kind, const int32_t box[6],
                   const int32_t stride[3], int32_t nfreq, const double *aux, double *acc);
/* one full step n (E half, then H half) in the reference order */
void orc_step(orc_grid *g, int32_t n);
/* run steps n0..n0+nsteps-1, sampling probes after each step:
 * probe p reads component pc[p] at (pi,pj,pk) into out[s*nprobe + p] */
void orc_run(orc_grid *g, int32_t n0, int32_t nsteps, int32_t nprobe,
             const int32_t *pc, const int32_t *pi, const int32_t *pj,
             const int32_t *pk, orc_real *out);

#ifdef __cplusplus
}
#endif
#endif
