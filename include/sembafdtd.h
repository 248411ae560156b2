/*
 * sembafdtd.h -- C ABI of the MI355X (gfx950) field time-stepping library.
 *
 * Drop-in for the hot path of semba-fdtd's time loop (ciclo_temporal,
 * src_main_pub/timestepping.F90:1248-1981).  The reference calls Fortran
 * module/internal subroutines on host arrays; this library keeps the fields,
 * media indices, CPML state and dispersive state resident in HBM and exposes
 * one entry point per reference call site.  A Fortran driver binds these
 * through iso_c_binding (semba_fdtd_amd/fortran/sembafdtd_gpu.F90).
 *
 * Conventions
 *  - Coordinates are total cells per axis (CPML slabs included); 0 is the
 *    outer face of the low slab.  nx, ny, nz are the reference's swept cell
 *    counts (fullsize, semba_fdtd.F90:1285-1309).
 *  - Host arrays of a field or media-index component are the reference's
 *    arrays as allocated (timestepping.F90:380-386, semba_fdtd.F90:1285-1300):
 *    the component's Alloc box [-1, e+1] on each axis, e being the last
 *    swept index (Ex: nx-1,ny,nz; Ey: nx,ny-1,nz; Ez: nx,ny,nz-1;
 *    Hx: nx,ny-1,nz-1; Hy: nx-1,ny,nz-1; Hz: nx-1,ny-1,nz), x fastest, so
 *    idx(i,j,k) = ((k+1)*EY + (j+1))*EX + (i+1) with EX = e_x+3, EY = e_y+3.
 *    A Fortran caller passes Ex(sgg%Alloc(iEx)%XI:...) etc. unchanged.
 *  - 1-D tables are indexed i+1 from -1: Idxe/Idye/Idze and the CPML
 *    profiles span Alloc of Hx/Hy/Hz (n+3 entries), Idxh/Idyh/Idzh span
 *    Alloc of Ex/Ey/Ez (n+2 entries).
 *  - Every function returns 0 on success or a negative FDTD_ERR_* code;
 *    fdtd_last_error() returns the message.  A handle is bound to one HIP
 *    stream and is not re-entrant; distinct handles may run concurrently.
 *  - The library never falls back to the CPU: without a usable gfx950 device
 *    fdtd_create fails with FDTD_ERR_DEVICE.
 */
#ifndef SEMBAFDTD_H
#define SEMBAFDTD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FDTD_ABI_VERSION 1

enum {
  FDTD_OK = 0,
  FDTD_ERR_ARG = -1,
  FDTD_ERR_DEVICE = -2,
  FDTD_ERR_NOMEM = -3,
  FDTD_ERR_STATE = -4,
  FDTD_ERR_HIP = -5
};

/* Arithmetic mode.
 * FAST : FMA contraction allowed, conjugate pole pairs folded into one
 *        complex64 state (fp32), the production path.
 * EXACT: no FMA contraction, complex(8) pole state updated pole by pole in
 *        the reference's order; bit-identical to the reference loops. */
enum { FDTD_MODE_FAST = 0, FDTD_MODE_EXACT = 1 };

/* Field storage in HBM.  FP16 keeps each field as binary16 times a per-array
 * scale with fp32 arithmetic in registers (CPML psi and pole state stay fp32);
 * FAST mode only; default scales 1 for E and 1/376.73 for H
 * (fdtd_set_field_scale).  The granular entry points need FP32 storage. */
enum { FDTD_STORE_FP32 = 0, FDTD_STORE_FP16 = 1 };

/* field components */
enum { FDTD_EX = 0, FDTD_EY, FDTD_EZ, FDTD_HX, FDTD_HY, FDTD_HZ };

/* media kind bits (per media index) */
enum { FDTD_KIND_PEC = 1, FDTD_KIND_PMC = 2, FDTD_KIND_EDISP = 4, FDTD_KIND_MDISP = 8 };

typedef struct fdtd_handle fdtd_handle;

typedef struct {
  int32_t nx, ny, nz; /* total swept cells per axis, CPML included */
  int32_t mode;       /* FDTD_MODE_* */
  int32_t storage;    /* FDTD_STORE_* */
  int32_t device;     /* HIP device ordinal */
  void *stream;       /* hipStream_t to run on; NULL = library-owned stream */
} fdtd_grid_desc;

/* ---- lifetime --------------------------------------------------------- */
/* Allocates the six fields and media matrices (launch_simulation's field
 * allocation, timestepping.F90:375-402); fields start at zero. */
int fdtd_create(const fdtd_grid_desc *desc, fdtd_handle **out);
/* Frees everything (Destroy_All_exceptSGGMxx, timestepping.F90:3177-3227).
 * FDTD_ERR_STATE, with nothing freed, while h is admitted to a batch. */
int fdtd_destroy(fdtd_handle *h);
/* NUL-terminated message of the last failure on h (or of fdtd_create when
 * h is NULL). */
const char *fdtd_last_error(const fdtd_handle *h);
int fdtd_abi_version(void);

/* ---- set-up ----------------------------------------------------------- */
/* Coefficient tables G1/G2/GM1/GM2 [0..nmedia] (calc_G1G2Gm1Gm2,
 * calc_constants.F90:36-148, after InitEDispersives overrides them) and media
 * kind bits [0..nmedia]. */
int fdtd_set_media(fdtd_handle *h, int32_t nmedia, const float *g1,
                   const float *g2, const float *gm1, const float *gm2,
                   const int32_t *kind);
/* Media index matrix of one component (sggMiEx..sggMiHz in their Alloc
 * shape; semba_fdtd.F90:93, built by read_geomData). */
int fdtd_set_media_index(fdtd_handle *h, int32_t comp, const int16_t *mi);
/* Inverse spacings Idxe,Idye,Idze,Idxh,Idyh,Idzh (timestepping.F90:314-370,
 * already stretched by the CPML, borderscpml.F90:1556-1724). */
int fdtd_set_spacing(fdtd_handle *h, const float *idxe, const float *idye,
                     const float *idze, const float *idxh, const float *idyh,
                     const float *idzh);
/* CPML (InitCPMLBorders, borderscpml.F90:90-473): layers[a*2+s] for axis
 * a = x,y,z and side s = low,high (0 = no CPML on that face); profiles
 * P_be, P_ce (E) and P_bm, P_cm (H) per axis, length n+3.  psi starts at 0. */
int fdtd_set_cpml(fdtd_handle *h, const int32_t layers[6],
                  const float *pbe_x, const float *pce_x, const float *pbm_x,
                  const float *pcm_x, const float *pbe_y, const float *pce_y,
                  const float *pbm_y, const float *pcm_y, const float *pbe_z,
                  const float *pce_z, const float *pbm_z, const float *pcm_z);
/* Electric pole-residue media (InitEDispersives, electricdispersive.F90:
 * 77-298): for each of ndisp media its index, pole count (<= 8) and the
 * complex(8) coefficients kappa_p, beta_p, G3_p as [ndisp][8][re,im].
 * Pole state starts at zero. */
int fdtd_set_edispersive(fdtd_handle *h, int32_t ndisp, const int32_t *medium,
                         const int32_t *npoles, const double *kappa,
                         const double *beta, const double *g3);
/* Magnetic pole-residue media (InitMdispersives, magneticdispersive.F90:
 * 77-290): the same recursion on H with mu11/Sigmam11, coefficients
 * kappa_p, beta_p, GM3_p as [ndisp][8][re,im].  Pole state starts at zero. */
int fdtd_set_mdispersive(fdtd_handle *h, int32_t ndisp, const int32_t *medium,
                         const int32_t *npoles, const double *kappa,
                         const double *beta, const double *gm3);
/* First-order Mur absorbing faces (InitMURBorders, bordersmur.F90:79-314):
 * faces[6] = x-, x+, y-, y+, z-, z+ (nonzero: absorbing), cab1 the
 * [6][nmedia+1] coefficients of fdtd_build_mur.  Call after fdtd_set_media.
 * The stored previous-step planes start at zero. */
int fdtd_set_mur(fdtd_handle *h, const int32_t faces[6], const float *cab1);
/* Time step and time vector (crea_timevector, timestepping.F90:3230-3240):
 * tiempo[0..nlast]; NULL builds tiempo(n) = tiempo(n-1) + dt from 0. */
int fdtd_set_time(fdtd_handle *h, double dt, int32_t nlast, const double *tiempo);
/* One nodal source (InitnodalSources/createnodal, nodalsources.F90:245-300):
 * component, hard (1) or soft (0), box {xi,xe,yi,ye,zi,ze} inside the sweep,
 * amplitude and excitation samples evol[0..numus] spaced deltaevol. */
int fdtd_add_nodal(fdtd_handle *h, int32_t comp, int32_t hard,
                   const int32_t box[6], float amp, float deltaevol,
                   int32_t numus, const float *evol);

/* Node coordinates of the discretisation lines (sgg%LineX/Y/Z,
 * preprocess_geom.F90:5330-5347,5672-5712), each over [-1, n+1] (n+3 entries);
 * the plane waves place their fields with them. */
int fdtd_set_lines(fdtd_handle *h, const float *line_x, const float *line_y,
                   const float *line_z);
/* One TF/SF plane-wave box (sgg%PlaneWave(j), fdetypes.F90:458-465, set up by
 * InitPlaneWave, planewaves.F90:72-770): box {esqx1,esqx2,esqy1,esqy2,esqz1,
 * esqz2} in total cells, nmodes (<= 4) modes with direction p[m][3], E
 * amplitude e[m][3] and start uncertainty incert[m], excitation samples
 * evol[0..numus] spaced deltaevol, and the real(4) eps0/mu0 of the run. */
int fdtd_add_planewave(fdtd_handle *h, const int32_t box[6], int32_t nmodes, const float *p,
                       const float *e, const float *incert, float deltaevol, int32_t numus,
                       const float *evol, float eps0, float mu0);

/* Work decomposition of the face passes: blocks of 64 x rows threads (rows =
 * 1, 2 or 4) that march zchunk z planes each, every lane owning xcells (1 or
 * 4) consecutive x cells of the fused passes.  No counterpart in the
 * reference (its loops are OpenMP over k); chosen by the autotuner
 * (semba_fdtd_amd.tuning) or left at the default by grid depth: 2 rows x 4
 * planes up to 160 planes (nz + 1), 4 x 8 up to 640, 4 x 16 beyond; 1 cell.
 * Either storage runs either.  fdtd_get_tiling returns the
 * values as set (a chunk longer than the grid runs as the whole grid). */
int fdtd_set_tiling(fdtd_handle *h, int32_t rows, int32_t zchunk, int32_t xcells);
int fdtd_get_tiling(const fdtd_handle *h, int32_t *rows, int32_t *zchunk, int32_t *xcells);

/* Average duration in ms, over reps launches on the grid's stream, of the
 * fused E pass with its pole update (ms[0]), the fused H pass (ms[1]) and the
 * pole update alone (ms[2]); sources and probes are not run.  The fields
 * advance while this runs: save the state first to keep it.  No counterpart
 * in the reference; used by the autotuner. */
int fdtd_time_halves(fdtd_handle *h, int32_t reps, float ms[3]);

/* One launch per step in fdtd_run: blocks of 64 x rows threads (rows 4, 8 or
 * 16; 0 turns it off) marching zchunk z planes compute E and then H of each
 * plane, reading the state of step n from one copy and writing step n+1 to a
 * second copy held for the purpose (fields, CPML psi and pole state: twice
 * the state's memory).  fdtd_run advances pairs of steps so that the state is
 * back in its usual place when it returns; an odd last step runs the two
 * passes.  The results equal the two-pass step bit for bit.  fp32 storage
 * only.  No counterpart in the reference. */
int fdtd_set_step_mode(fdtd_handle *h, int32_t rows, int32_t zchunk);
int fdtd_get_step_mode(const fdtd_handle *h, int32_t *rows, int32_t *zchunk);
/* Tile classes of the one-launch step: 0-7 carry bit 0 for tiles that touch a
 * CPML slab, bit 1 for tiles that meet a pole-state box, bit 2 for tiles that
 * touch a TF/SF face or a nodal cell (one lane per x cell); class 8 holds the
 * plain tiles run four x cells per lane; classes 9-16 are classes 0-7 on
 * half-width tiles (32 lanes per row, two rows per wave) for x pieces of at
 * most 31 columns, such as the x CPML slabs. */
#define FDTD_STEP_CLASSES 17
/* Rows per block (4, 6, 8 or 16; 0: the default) of the step kernels of the
 * tile classes in the bit mask classes (bit c: class c).  A class on fewer
 * rows than the step runs the step's tiles cut into bands of rows - 1 rows;
 * for classes 0-7 rows - 1 must divide the step's rows - 1 (16-row steps: 4
 * or 6), class 8 runs 4, 8 or 16 rows; other values leave the class on its
 * default.  Classes 9-16 run blocks of that many waves (any of the four
 * values up to the step's rows), each tile owning up to 2 rows - 1 rows.
 * Smaller blocks give a kernel that needs many registers per lane more of
 * them; by default, under 16-row steps, the CPML-only class and the class
 * with CPML, poles and sources run 4-row blocks, the half-width CPML +
 * sources tiles 8 waves and the half-width CPML + poles + sources tiles 4
 * (at most the step's rows), every other class the step's rows.  fdtd_get_step_class_rows gives the rows each class runs with.
 * The results do not change.  No counterpart in the reference; used by the
 * autotuner. */
int fdtd_set_step_class_rows(fdtd_handle *h, int32_t classes, int32_t rows);
int fdtd_get_step_class_rows(const fdtd_handle *h, int32_t rows[FDTD_STEP_CLASSES]);
/* Plain tiles of the one-launch step run four x cells per lane (on = 1, the
 * default) or one (on = 0).  The results do not change.  No counterpart in
 * the reference. */
int fdtd_set_step_vector(fdtd_handle *h, int32_t on);
/* Average ms of one one-launch step (the step kernels alone), over reps
 * pairs; the fields advance. */
int fdtd_time_step(fdtd_handle *h, int32_t reps, float *ms);
/* The tile classes of the one-launch step (FDTD_STEP_CLASSES), one kernel
 * launch each: cls[q] the class, cells[q] the Yee cells (nx*ny*nz grid) its
 * tiles own, bytes[q] the compulsory HBM bytes of one launch (every field
 * read and written once and every media index read once per owned cell, 8 B
 * per CPML psi value, per pole-state cell 8 B of previous field and 16 B per
 * stored complex64 pole state in FAST, 32 B per complex128 one in EXACT).
 * With reps > 0, ms[q] is the average duration of one launch of the class's
 * kernel over reps launches of the first half of a step pair (the state of
 * step n is read and the second copy rewritten; the grid does not advance).
 * Returns the number of classes (<= FDTD_STEP_CLASSES) or a negative code.  No counterpart in
 * the reference; used by the benchmark. */
int fdtd_step_classes(fdtd_handle *h, int32_t reps, int32_t cls[FDTD_STEP_CLASSES],
                      double cells[FDTD_STEP_CLASSES], double bytes[FDTD_STEP_CLASSES],
                      float ms[FDTD_STEP_CLASSES]);

/* ---- the hot path ----------------------------------------------------- */
/* Fused E half step: Advance_Ex/Ey/Ez (timestepping.F90:1286-1297) +
 * AdvanceelectricCPML (:1381) in one HBM pass, AdvanceEDispersiveE (:1416)
 * over the dispersive edges, then AdvancePlaneWaveE (:1425) and AdvanceNodalE
 * (:1441) at tiempo(n). */
int fdtd_advance_e(fdtd_handle *h, int32_t n);
/* Fused H half step: Advance_Hx/Hy/Hz (:1497-1507) + AdvanceMagneticCPML
 * (:1535) in one pass, then AdvancePlaneWaveH (:1598) and AdvanceNodalH
 * (:1615) at tiempo(n)+dt/2. */
int fdtd_advance_h(fdtd_handle *h, int32_t n);
/* The same pieces one call site at a time, for drivers that interleave other
 * host modules between them. */
int fdtd_advance_e_yee(fdtd_handle *h);        /* Advance_Ex/Ey/Ez */
int fdtd_advance_e_cpml(fdtd_handle *h);       /* AdvanceelectricCPML */
int fdtd_advance_e_dispersive(fdtd_handle *h); /* AdvanceEDispersiveE */
int fdtd_advance_planewave_e(fdtd_handle *h, int32_t n); /* AdvancePlaneWaveE */
int fdtd_advance_nodal_e(fdtd_handle *h, int32_t n); /* AdvanceNodalE */
int fdtd_advance_h_yee(fdtd_handle *h);        /* Advance_Hx/Hy/Hz */
int fdtd_advance_h_cpml(fdtd_handle *h);       /* AdvanceMagneticCPML */
int fdtd_advance_h_dispersive(fdtd_handle *h); /* AdvanceMdispersiveH */
int fdtd_advance_planewave_h(fdtd_handle *h, int32_t n); /* AdvancePlaneWaveH */
int fdtd_advance_nodal_h(fdtd_handle *h, int32_t n); /* AdvanceNodalH */
int fdtd_advance_mur(fdtd_handle *h);          /* AdvanceMagneticMUR */
/* nsteps full steps n0 .. n0+nsteps-1 (E half then H half, probes sampled and
 * frequency-domain sums advanced after each step as UpdateObservation does,
 * timestepping.F90:1725-1742), replayed from a captured HIP graph.
 * n0 >= 0.  Every step samples the point probes once: a run that would take
 * more samples than the probe buffer has room for since the last
 * fdtd_read_probes is refused (FDTD_ERR_ARG) before any launch.
 * Stepping calls (fdtd_run, fdtd_advance_*, fdtd_update_observation,
 * fdtd_sample_probes, the timing calls) return FDTD_ERR_STATE while h is
 * admitted to a batch: the batch advances it. */
int fdtd_run(fdtd_handle *h, int32_t n0, int32_t nsteps);
/* Steps per replay of fdtd_run's captured two-pass step (1..64; default
 * 16): one graph records that many steps back to back and is replayed while
 * that many steps remain, the rest one step per replay.  Fewer replays cost
 * less launch time on small grids (64^3 cavity: 76.7 -> 72.2 us per step
 * from 1 to 16); the results do not depend on it.  No counterpart in the
 * reference. */
int fdtd_set_graph_steps(fdtd_handle *h, int32_t steps);

/* ---- observation ------------------------------------------------------ */
/* Point probes (observation.F90:2620-2672): component and Yee index each;
 * samples are buffered on the device, capacity = bufsteps samples. */
int fdtd_set_probes(fdtd_handle *h, int32_t nprobe, const int32_t *comp,
                    const int32_t *i, const int32_t *j, const int32_t *k,
                    int32_t bufsteps);
/* Copies the buffered samples [nsamples][nprobe] to out (capacity
 * max_samples rows), clears the buffer, returns the count in *nsamples.
 * Calls that sample the probes (fdtd_run, fdtd_update_observation,
 * fdtd_sample_probes, fdtd_batch_step) refuse to take more than bufsteps
 * samples between two reads. */
int fdtd_read_probes(fdtd_handle *h, float *out, int32_t max_samples,
                     int32_t *nsamples);
/* Sample probes once now (used by drivers that step with fdtd_advance_*). */
int fdtd_sample_probes(fdtd_handle *h);
/* Frequency-domain observation box (freq-domain probes, observation.F90:
 * 3527-3600): kind 0 sums Ex,Ey,Ez, kind 1 Hx,Hy,Hz at the points of box
 * {i1,i2,j1,j2,k1,k2} whose coordinates are multiples of stride[3], for nfreq
 * real(4) frequencies [Hz]: complex(8) running sums of dt*F(t_n)*exp(-j2 pi f
 * t_n) (H at t_n + dt/2).  *id names the box. */
int fdtd_add_dft(fdtd_handle *h, int32_t kind, const int32_t box[6], const int32_t stride[3],
                 int32_t nfreq, const float *freq, int32_t *id);
/* Complex values held by box id (points x 3 x nfreq), -1 for a bad id. */
int64_t fdtd_dft_size(const fdtd_handle *h, int32_t id);
/* The running sums of box id as [k'][j'][i'][comp][freq][re,im] doubles
 * (valor3DComplex's order); capacity counts complex values. */
int fdtd_read_dft(fdtd_handle *h, int32_t id, double *out, int64_t capacity);
/* UpdateObservation at step n (timestepping.F90:1727): sample the point probes
 * and add step n to the frequency-domain sums.  fdtd_run does this itself. */
int fdtd_update_observation(fdtd_handle *h, int32_t n);

/* ---- field access ----------------------------------------------------- */
int fdtd_get_field(fdtd_handle *h, int32_t comp, float *host_alloc);
int fdtd_set_field(fdtd_handle *h, int32_t comp, const float *host_alloc);
/* fdtd_get_field without any device allocation, for callers that cannot
 * allocate (loops that must not call hipMalloc, harnesses that forbid it).
 * No counterpart in the reference.  It stages runs of z planes through a
 * fixed 16 MiB device array per GPU (held until the process ends) instead of
 * a buffer it allocates; a component whose z plane exceeds 16 MiB as fp32 is
 * refused with FDTD_ERR_ARG.  Same output as fdtd_get_field.  All calls on one
 * GPU share that array: calls that overlap on one GPU race. */
int fdtd_get_field_noalloc(fdtd_handle *h, int32_t comp, float *host_alloc);
/* Scale of a binary16 field array (value = stored * scale); stored values keep
 * their physical value.  No effect on FP32 storage. */
int fdtd_set_field_scale(fdtd_handle *h, int32_t comp, float scale);
int fdtd_synchronize(fdtd_handle *h);
/* Checkpoint / resume (StoreFields and the modules' restart files,
 * resuming.F90:394-496): the whole time-stepping state -- fields, CPML psi,
 * pole state, frequency-domain sums -- with the step n to resume at, as one
 * opaque byte string of fdtd_state_size bytes.  A state loads into a grid with
 * the same set-up and mode; *n receives the saved step. */
int64_t fdtd_state_size(fdtd_handle *h);
int fdtd_save_state(fdtd_handle *h, int32_t n, void *buf, int64_t size);
int fdtd_load_state(fdtd_handle *h, const void *buf, int64_t size, int32_t *n);
/* Device bytes held by the handle. */
int64_t fdtd_device_bytes(const fdtd_handle *h);
/* Free and total bytes of device memory (hipMemGetInfo on `device`), for
 * callers that admit grids by the memory the device has left.  No
 * counterpart in the reference (it runs one grid per process). */
int fdtd_device_memory(int32_t device, int64_t *free_bytes, int64_t *total_bytes);
/* Number of dispersive edges of component comp (0..2 E, 3..5 H) holding pole
 * state (the node lists of InitEDispersives, electricdispersive.F90:148-244,
 * and InitMdispersives, magneticdispersive.F90:139-265); valid after the
 * first advance/run call, -1 before. */
int64_t fdtd_dispersive_edges(const fdtd_handle *h, int32_t comp);

/* ---- host-side coefficient tables (built once, uploaded with the set-up
 * calls above).  They evaluate the reference's expressions in its kinds. */
/* dt: semba_fdtd.F90:1211-1250 (forcecfl = the -cfl switch). */
double fdtd_courant_dt(float dxmin, float dymin, float dzmin, float eps0, float mu0,
                       float cfl, int32_t forcecfl, double dt_in);
/* calc_G1G2Gm1Gm2, calc_constants.F90:36-148; arrays [0..nmedia]. */
int fdtd_build_g_tables(int32_t nmedia, const float *epr, const float *sigma,
                        const float *mur, const float *sigmam, const int32_t *kind,
                        double dt, float eps0, float mu0, float *g1, float *g2,
                        float *gm1, float *gm2);
/* spacings, timestepping.F90:314-370: D[-2..n+1] -> de, dh, 1/de, 1/dh on
 * [-1..n+1] (length n+3). */
int fdtd_build_spacing(int32_t n, const float *D, float *de, float *dh, float *ide,
                       float *idh);
/* discretisation lines of one axis: D[-2..n+1] -> line[-1..n+1], origin 0
 * at the first interior node Llo (preprocess_geom.F90:5330-5347,5690-5712). */
int fdtd_build_lines(int32_t n, int32_t Llo, int32_t Lhi, const float *D, float *line);
/* CPML profiles of one axis, calc_cpmlconstants (borderscpml.F90:1465-1733);
 * stretches ide/idh in place. */
int fdtd_build_cpml_axis(int32_t n, int32_t Llo, int32_t Lhi, float orden_lo,
                         float orden_hi, float R_lo, float R_hi, float alphamaxpar,
                         float alphaorden, float kappamaxpar, double dt, float eps0,
                         float mu0, const float *de, const float *dh, float *pbe,
                         float *pce, float *pbm, float *pcm, float *ide, float *idh);
/* pole-residue coefficients and the medium's G1/G2 (electricdispersive.F90:
 * 98-146); c, a, kappa, beta, g3 are [npoles][re,im]. */
int fdtd_build_ade(int32_t npoles, const double *c, const double *a, float eps11,
                   float sigma11, double dt, double *kappa, double *beta, double *g3,
                   float *g1, float *g2);
/* First-order Mur coefficients CAB1 (calc_murconstants, bordersmur.F90:
 * 316-363) for the six faces x-, x+, y-, y+, z-, z+: idh_face holds the dual
 * inverse spacing at the face's end of the Alloc range (Idxh(-1), Idxh(nx),
 * Idyh(-1), Idyh(ny), Idzh(-1), Idzh(nz), after any CPML stretching);
 * cab1 is [6][nmedia+1]. */
int fdtd_build_mur(int32_t nmedia, const float *epr, const float *mur, double dt, float eps0,
                   float mu0, const float idh_face[6], float *cab1);
/* evolucion, nodalsources.F90:317-356. */
float fdtd_evolucion(float t, float deltaevol, int32_t numus, const float *evol);

/* ---- batches of independent grids ---------------------------------------
 * No counterpart in the reference (one solve per process there).  A batch
 * advances many grids, each set up as above, together: one persistent launch
 * per half step covers the face passes of every grid, the rest of the step
 * is grouped across grids, and the step is replayed from a HIP graph.  Each
 * grid keeps its own step counter, so grids join (admit) and leave (retire)
 * between steps: continuous batching.  Results equal the grid advanced
 * alone. */
typedef struct fdtd_batch fdtd_batch;
int fdtd_batch_create(int32_t device, void *stream, fdtd_batch **out);
int fdtd_batch_destroy(fdtd_batch *b);
const char *fdtd_batch_last_error(const fdtd_batch *b);
/* Admit grid h at its step n0; *slot names it in the batch.  While admitted
 * its own calls (reads, set-up) run on the batch's stream, and its stepping
 * calls and fdtd_destroy are refused (FDTD_ERR_STATE) until it is retired
 * (or the batch destroyed).  A grid belongs to at most one batch: admitting
 * it again, to this batch or another, is refused (FDTD_ERR_STATE). */
int fdtd_batch_admit(fdtd_batch *b, fdtd_handle *h, int32_t n0, int32_t *slot);
int fdtd_batch_retire(fdtd_batch *b, int32_t slot);
/* Step the grid of slot will run next (-1 for an empty slot). */
int64_t fdtd_batch_step_of(const fdtd_batch *b, int32_t slot);
/* Tiling of the batched face passes (rows 1, 2 or 4; zchunk planes; xcells 1:
 * one cell per lane, each block taking a run of 64 * rows consecutive columns
 * of a grid's x-y plane, x fastest, so that narrow grids leave no lane idle;
 * xcells 4: four x cells per lane over 256-cell x segments) and the
 * persistent launch's block count (0: 16384).  Either storage runs either. */
int fdtd_batch_set_tiling(fdtd_batch *b, int32_t rows, int32_t zchunk, int32_t xcells,
                          int32_t blocks);
/* One launch per step for the batch (rows 4, 8 or 16; 0: the face passes
 * above): every grid holds a second copy of its state and its own tile plan
 * (as fdtd_set_step_mode, which this sets on each admitted grid), the batch
 * advances pairs of steps with one launch per tile class over the tiles of
 * all grids, and an odd last step runs the face passes.  fp32 storage only.
 * The results equal the grids' own steps bit for bit.  No counterpart in the
 * reference. */
int fdtd_batch_set_step_mode(fdtd_batch *b, int32_t rows, int32_t zchunk);
/* nsteps steps of every admitted grid. */
int fdtd_batch_step(fdtd_batch *b, int32_t nsteps);
int fdtd_batch_synchronize(fdtd_batch *b);

/* ---- device-side scene building for grids too large to build on the host.
 * Same rules as the reference preprocessing, applied in HBM. */
int fdtd_media_fill(fdtd_handle *h, int16_t medium);
/* PEC outer walls (healer.F90:1442-1617): faces bit mask xlo,xhi,ylo,yhi,zlo,zhi. */
int fdtd_media_pec_walls(fdtd_handle *h, int32_t faces);
/* Voxelised sphere (cells whose centre is inside) painted with
 * CreateVolumeMM's priority rule (healer.F90:66-234): an edge takes the medium
 * when prio[medium] > prio[current], a face when >=. prio has nmedia+1
 * entries. */
int fdtd_media_paint_sphere(fdtd_handle *h, float cx, float cy, float cz,
                            float radius, int16_t medium, int32_t nprio,
                            const int32_t *prio);
/* Axis-aligned box of cells [x0,x1)x[y0,y1)x[z0,z1) painted the same way. */
int fdtd_media_paint_box(fdtd_handle *h, int32_t x0, int32_t x1, int32_t y0,
                         int32_t y1, int32_t z0, int32_t z1, int16_t medium,
                         int32_t nprio, const int32_t *prio);

#ifdef __cplusplus
}
#endif
#endif
