// fdtd_internal.h -- device data layout shared by the host code and the
// gfx950 kernels.
//
// HBM layout of one grid
//   fields      6 arrays (Ex Ey Ez Hx Hy Hz), each (nz+3) planes x (ny+3) rows
//               x px elements; cell (i,j,k) at ((k+1)*(ny+3)+(j+1))*px + xoff+i.
//               Grids at least 256 cells wide: xoff = 32 puts x = 0 on a
//               128-byte line boundary in every row (px a multiple of 32).
//               Narrower grids: xoff = 4 and px a multiple of 4 (16 bytes),
//               so that short rows do not drag whole padding lines along.
//               Either way px leaves room for x up to n+7, so the last
//               16-byte vector of a row and its right neighbour stay in the
//               row, and the Alloc padding planes/rows/columns (-1 and n+1)
//               stay zero forever.
//   media idx   6 int16 arrays with the same geometry as the fields.
//   tables      G1/G2/GM1/GM2 and kind bits per media index; inverse spacings
//               (length n+3, index i+1); CPML profiles per axis.
//   psi         12 compact slab arrays (component, stretched axis): along the
//               stretched axis only the low and high slab layers are stored
//               (Llo + Lhi entries), the two other axes span the component's
//               sweep; x-fastest.
//   pole state  per E component (electric media) and per H component
//               (magnetic media) a dense x-fastest box spanning the
//               component's dispersive edges (Eprev and one complex state per
//               stored pole); an edge finds its slot from its coordinates, so
//               the state loads issue alongside the field loads.  Only
//               dispersive edges touch it, so a sparse box costs capacity,
//               not bandwidth.
#pragma once
#include <stdint.h>
#ifdef __HIPCC__
#include <hip/hip_fp16.h>
#endif

namespace sfd {

constexpr int kPMax = 8;     // poles per dispersive medium
constexpr int kWave = 64;    // CDNA wavefront
constexpr int kXOff = 32;    // x = 0 offset inside a row of a wide grid (elements)
constexpr int kXOffNarrow = 4;   // the same for grids under 256 cells wide

enum { PSI_EXY = 0, PSI_EXZ, PSI_EYX, PSI_EYZ, PSI_EZX, PSI_EZY,
       PSI_HXY, PSI_HXZ, PSI_HYX, PSI_HYZ, PSI_HZX, PSI_HZY };

// Pole-residue state of one field: the electric media acting on E
// (electricdispersive.F90) or the magnetic media acting on H
// (magneticdispersive.F90); the two share one recursion.
struct AdeSet {
  int has;                      // some edge of the field is dispersive
  int ndisp;                    // dispersive media (slots)
  const int32_t *dslot;         // [nmedia+1] dispersive slot or -1
  int abox[3][6];               // per component: x0,y0,z0,bx,by,bz of the
                                // box holding its dispersive edges
  float *prev[3];               // field at the previous step, per box slot
  void *jst[3];                 // FAST: float2[P][N], EXACT: double2[P][N]
  long long nstate[3];
  int maxp;                     // pole planes of each state array (>= 1)
  const int32_t *anp;           // per slot: stored poles
  const float *adef;            // FAST: per slot per pole {kr,ki,bdtr,bdti,g3r,g3i}
  const double *aded;           // EXACT: per slot per pole {kr,ki,br,bi,g3r,g3i}
};

struct DevGrid {
  int nx, ny, nz;
  int nmedia;
  long long px, pxy;            // row pitch and plane stride in elements
  int xoff;                     // x = 0 offset inside a row (kXOff or kXOffNarrow)
  float *f[6];                  // fp32 fields, or binary16 ones when fp16
  int fp16;                     // fields stored as binary16 * fs[c]
  float fs[6], fi[6];           // per-array scale and its inverse
  const int16_t *mi[6];
  const float *g1, *g2, *gm1, *gm2;
  const float *id[6];           // Idxe Idye Idze Idxh Idyh Idzh
  // CPML
  int L[3][2];
  const float *pbe[3], *pce[3], *pbm[3], *pcm[3];
  float *psi[12];
  int has_cpml;
  AdeSet ade[2];                // 0: electric media (E), 1: magnetic media (H)
  double dt;
};

// blocks of (64 x ty) threads march z in chunks of kc planes.  nblk tiles
// are numbered x fastest; remap spreads contiguous tile ranges over the 8
// XCDs so that neighbouring tiles share an L2.
struct TileMap {
  int tx, ty, tz;   // tiles per axis
  int kc;           // planes per z chunk
  int nblk;
  int remap;
  int bdy;          // rows per block (block = 64 x bdy threads, <= 256)
  int vec;          // x cells per lane of the fused passes (1 or 4)
  int flat;         // 1: tiles are runs of 64 * bdy columns of the x-y plane
                    // (one cell per lane), tx per plane, ty = 1
};

// First-order Mur (bordersmur.F90): one tangential H component on one
// absorbing face.  The plane just outside the component's sweep (ghost) is
// advanced from the first plane inside (inner); past keeps both planes of
// the previous step.
struct MurRun {
  int comp;                // 3..5
  int axis, ghost, inner;  // face normal and the two planes along it
  int ua, va;              // the two other axes (u fastest)
  int nu, nv;              // extents: the component's sweep along ua, va
  long long start;         // first cell of this run in the launch
  float *past;             // [2][nv][nu]: ghost plane, then inner plane
  const float *cab1;       // [nmedia+1] coefficients of the face
};

struct NodalCell {   // one cell touched by a nodal source, in reference order
  long long p;       // field offset
  int comp;
  int hard;
  int src;           // source index (value table column)
  int skip;          // media is PEC (E) / PMC (H): leave untouched
  float coef;        // soft: G2*Ia*Ib (already rounded in reference order)
  float amp;
};

// TF/SF plane waves (planewaves.F90).  PwWave: the constants InitPlaneWave
// derives for one wave; PwFace: one face of one component, as one loop of
// AdvancePlaneWaveE/H applies it.
constexpr int kPwModes = 4;

struct PwWave {
  float pn[kPwModes][3];      // normalised directions
  float fpw[6][kPwModes];     // incident amplitudes of Ex..Hz per mode
  float dist0[kPwModes];      // distanciaInicial
  int nmodes;
  int numus;
  float deltaevol;
  float cluz;                 // 1/sqrt(eps0 mu0) in real(4)
  const float *evol;          // device, numus+1 samples
};

struct PwFace {
  int comp, axis, fixed;      // updated component, face normal, face index
  int lo[3], hi[3];           // ranges along the two free axes
  int nf, off;                // incident component, its offset along axis
  float sign, id;             // +1/-1 and the inverse spacing at `fixed`
  int wave;
  long long inc;              // one-launch steps: the face's first cell in the
                              // incident-field buffer (cells ua fastest)
};

// One TF/SF face of one grid with everything its launch needs, so that the
// faces of one group (one wave, one face axis) of many grids run in a single
// launch: faces of a group never share a cell within a grid.
struct PwRun {
  PwFace f;
  float *F;                   // the updated component
  int fp16;                   // F holds binary16 values * fs
  float fs, fi;
  long long px, pxy;          // its row pitch and plane stride
  int xoff;                   // x = 0 offset inside a row
  const PwWave *wave;
  const float *phys;
  int pstride;
  int half;
  const double *tiempo;
  const int *dn;              // device step counter (or null: the launch's n)
  double dt;
  float G;
  int nu;                     // cells along the first free axis
};

// Frequency-domain observation box (observation.F90:3527-3600): complex(8)
// running sums acc[pt][comp][freq] += aux[n][freq] * F_comp(pt), the points
// being the stride multiples of the box (first[a] + t*stride[a], cnt[a] of them).
struct DftDev {
  int kind;                   // 0: Ex,Ey,Ez  1: Hx,Hy,Hz
  int first[3], stride[3], cnt[3];
  int nfreq;
  const double2 *aux;         // [step][freq]
  double2 *acc;               // [pt][3][freq], pt = (kt*cnt1 + jt)*cnt0 + it
};

struct ProbeDev {
  long long p;
  int comp;
};

// TF/SF faces per half that a one-launch step handles (one bit each in the
// lane and plane masks)
constexpr int kStepFaces = 64;

// One full step of one grid in a single z-marching launch (E then H of each
// plane, both fields read from gi and written to go; see k_step in
// fdtd_kernels.hip).  Tiles of 64 x R threads overlap their neighbours by one
// column and one row, and each z chunk recomputes the E plane above it, so
// every H value a tile writes has its E neighbours computed in the same
// block.  The plane-wave faces and nodal cells are applied inline, in the
// reference's order.
// the plane-wave faces and nodal cells of a grid, in device memory (read by
// the few cells they touch)
struct StepSrc {
  // TF/SF faces of the E and H halves, in the reference's face order
  int npw[2];
  const PwFace *pwf[2];
  const PwWave *waves;
  const float *phys;
  int pstride;
  const double *tiempo;
  double dt;
  float pwG[2];               // G2(1), GM2(1)
  // nodal cells of the E and H halves, in the reference's order
  int ncell[2];
  const NodalCell *cells[2];
  const float *values;        // [step][source]
  int nsrc;
};

// the array bases of one step direction, in device memory: the march reads
// them once per plane through a constant-space pointer instead of holding
// them in scalar registers across the loop (fields, media, CPML psi, pole
// state)
struct StepAux {
  const float *fi[6];         // fields of step n
  float *fo[6];               // fields of step n+1
  const int16_t *mi[6];       // media indices
  const uint8_t *mi8[6];      // the same in 8 bits (grids with at most 256 media)
  const float *psi_i[12];     // read copy
  float *psi_o[12];           // written copy
  const float *pbe[3], *pce[3], *pbm[3], *pcm[3];
  AdeSet ai[2], ao[2];        // pole state: read and written copies, E and H
  int L[3][2];
  int nx, ny, nz;
  double dt;
};

// One tile of a one-launch step: owned columns [x0, x0 + wx), rows
// [y0, y0 + wy) and planes [k0, k1).  The block also computes E on column
// x0 + wx, row y0 + wy and plane k1 (the halo: the E neighbours of its last H
// column, row and plane), and stores nothing there.
//   x: x0 | wx << 20    y: y0 | wy << 20    z: k0    w: k1
struct StepTile {
  int x, y, k0, k1;
};
// k1 holds the end plane in its low 20 bits; a four-cell tile whose cells
// all have one medium (every component, halo included) carries that medium
// index above them
constexpr unsigned kTileK1Mask = 0xfffffu;
constexpr int kTileMediumShift = 20;
#ifdef __HIPCC__
__host__ __device__
#endif
inline StepTile step_tile(int x0, int wx, int y0, int wy, int k0, int k1) {
  return StepTile{x0 | (wx << 20), y0 | (wy << 20), k0, k1};
}

// the incident-field pre-pass of one grid of a batch (launch_pw_incident's
// arguments)
struct PwIncJob {
  const PwFace *faces;
  int nf0, nf;
  const StepSrc *src;
  const int *dn;
  float *inc;
};

// a tile of one grid of a batch stepped in one launch with other grids' tiles
struct GridTile {
  StepTile t;
  int grid;                   // index into the launch's StepDev array
  int pad[3];
};

struct StepDev {
  DevGrid gi, go;             // read (step n) and write (step n+1) views
  const StepAux *aux;
  TileMap tm;                 // kc: planes per chunk of the tiles with sources
                              // (their LDS face masks span kc + 1 planes);
                              // nblk: the tiles of this launch
  const StepTile *tiles;      // the tiles of this launch's class
  int prev_from_field;        // tiles without sources take the pole update's
                              // previous field from the field of step n
  int lazy_prev;              // ... and then leave the stored previous field
                              // to a settling pass (launch_prev_settle)
  int narrow;                 // media read from the 8-bit copies mi8
  const uint8_t *mi8[6];
  // TF/SF faces per half (staged in LDS by the blocks that meet them), and
  // the bounding box of each half's nodal cells
  int npw[2];
  int ncell[2];
  int cellbox[2][6];
  const StepSrc *src;
  const int *dn;              // device step counter
  const float *inc;           // incident field of every face cell this step
  float pwG[2];               // G2(1), GM2(1)
};

// a batch of independent grids for the persistent face-pass launch: tile t of
// the batch belongs to the grid b with start[b] <= t < start[b+1]
struct BatchDev {
  const DevGrid *grids;       // device copies of the grids' descriptors
  const TileMap *tms;
  const int *start;           // [ngrid+1] prefix of tile counts
  int ngrid;
  int total;
  const int *tmed;            // per tile: the medium all its cells hold, or -1
                              // (null: every tile loads its media)
};

#ifdef __HIPCC__
// one field value of component c at p, whatever the storage
__device__ __forceinline__ float field_ld(const DevGrid &g, int c, long long p) {
  return g.fp16 ? __half2float(reinterpret_cast<const __half *>(g.f[c])[p]) * g.fs[c]
                : g.f[c][p];
}
__device__ __forceinline__ void field_st(const DevGrid &g, int c, long long p, float v) {
  if (g.fp16) reinterpret_cast<__half *>(g.f[c])[p] = __float2half_rn(v * g.fi[c]);
  else g.f[c][p] = v;
}

// The TF/SF and nodal arithmetic, shared by their own kernels and the
// one-launch step.  Written without contraction in every build, so both give
// the same bits.
// evolucion of the plane wave (planewaves.F90:824-851): linear interpolation
// of the excitation at the retarded time t - d/c, zero before the second
// sample and after the table.
__device__ __forceinline__ float pw_evol(const PwWave &w, float t, float d) {
#pragma clang fp contract(off)
  float ev = 0.0f;
  const long long nprev = (long long)(int)((t - d / w.cluz) / w.deltaevol);
  if (nprev + 1 <= w.numus && nprev > 0)
    ev = (w.evol[nprev + 1] - w.evol[nprev]) / w.deltaevol *
             ((t - d / w.cluz) - (float)nprev * w.deltaevol) + w.evol[nprev];
  return ev;
}

// Incid (planewaves.F90:778-853) of face f at cell cc and time tm: the
// incident component f.nf at the neighbouring position, summed over modes
__device__ __forceinline__ float pw_incid(const PwFace &f, const PwWave *waves,
                                          const float *phys, long long pstride, float tm,
                                          const int cc[3]) {
#pragma clang fp contract(off)
  int q[3] = {cc[0], cc[1], cc[2]};
  q[f.axis] += f.off;
  const float xf = phys[(size_t)(f.nf * 3 + 0) * pstride + q[0] + 1];
  const float yf = phys[(size_t)(f.nf * 3 + 1) * pstride + q[1] + 1];
  const float zf = phys[(size_t)(f.nf * 3 + 2) * pstride + q[2] + 1];
  const PwWave &w = waves[f.wave];
  float ehi = 0.0f;
  for (int m = 0; m < w.nmodes; ++m) {
    const float d = (xf * w.pn[m][0] + yf * w.pn[m][1] + zf * w.pn[m][2]) - w.dist0[m];
    ehi = ehi + w.fpw[f.nf][m] * pw_evol(w, tm, d);
  }
  return ehi;
}

// F +- G * Incid * Id (AdvancePlaneWaveE/H)
__device__ __forceinline__ float pw_apply(float v, float sign, float G, float ehi, float id) {
#pragma clang fp contract(off)
  return sign > 0.0f ? v + G * ehi * id : v - G * ehi * id;
}

// AdvanceNodalE/H on one cell: hard F = amp*f, soft F = F - coef*amp*f
__device__ __forceinline__ float nodal_apply(float v, const NodalCell &c, float f) {
#pragma clang fp contract(off)
  return c.hard ? c.amp * f : v - c.coef * c.amp * f;
}
#endif

// launchers, one set per arithmetic mode.  tmed (null: none) holds, per
// tile of the one-cell map tm, the medium of every cell of the tile or -1,
// as launch_pass_tile_media writes it; the one-cell passes skip the media
// loads of the tiles it names.
#define SFD_DECLARE_LAUNCHERS(NS)                                              \
  namespace NS {                                                             \
  int launch_e(const DevGrid &g, const TileMap &tm, const TileMap &tm4,      \
               const int *tmed, int yee, int cpml, int ade, void *stream);   \
  int launch_h(const DevGrid &g, const TileMap &tm, const TileMap &tm4,      \
               const int *tmed, int yee, int cpml, int ade, void *stream);   \
  int launch_pass_tile_media(const DevGrid &g, const TileMap &tm, int *out,  \
                             void *stream);                                  \
  size_t step_lds_bytes(const DevGrid &g, int rows, int npw, int kc);       \
  int launch_step(const StepDev &sd, int rows, int cls, int nuni, void *stream); \
  int launch_step_many(const StepDev *sds, const GridTile *tiles, int n, int narrow, int rows, \
                       int cls, size_t lds, void *stream);                   \
  size_t step4_lds_bytes(const DevGrid &g, int rows);                       \
  int launch_step4(const StepDev &sd, int rows, void *stream);              \
  size_t ade_lds_of(const DevGrid &g, int side);                            \
  int launch_ade_many(const DevGrid *grids, const int *list, int n, int side, \
                      size_t lds, int tbl, long long nmax, void *stream);     \
  int launch_pass_batch(const BatchDev &bd, int half, int blocks, int rows,  \
                        int vec, size_t lds, int fp16, void *stream);        \
  }
SFD_DECLARE_LAUNCHERS(fast)
SFD_DECLARE_LAUNCHERS(exact)

// mode-independent helper kernels (fdtd_aux.hip)
int launch_nodal(const DevGrid &g, const NodalCell *cells, int ncells,
                 const float *values, int nsrc, int n, const int *dn, int serial,
                 void *stream);
int launch_tick(int *dn, void *stream);
int launch_dft(const DevGrid &g, const DftDev &d, int n, const int *dn, void *stream);
// a group of TF/SF faces at step n (or each run's *dn): F += sign * G *
// Incid(t) * id, t the real(4) time stamp of the half step (E: tiempo(n), H:
// tiempo(n) + dt/2); runs[r] covers cells start[r] .. start[r+1]-1 of total
int launch_tfsf_list(const PwRun *runs, const long long *start, int nruns, long long total,
                     int n, void *stream);
// an 8-bit copy of n media indices, all below 256
int launch_media_narrow(const int16_t *src, uint8_t *dst, long long n, void *stream);
// a dense box of es-byte elements (es 2 or 4), ext = its x, y, z extents,
// moved between a packed device buffer and a grid array
int launch_box_move(void *grid, void *packed, size_t es, const long long ext[3], long long px,
                    long long pxy, int xoff, int to_grid, void *stream);
// the same box read out of a field array (es 4: fp32, es 2: binary16 times
// scale) into a packed fp32 buffer
// the stored previous field of every dispersive cell a step tile owns set
// to the cell's field (the tiles of the pole classes without sources, after
// steps that left it to this pass)
int launch_prev_settle(const DevGrid &g, const StepTile *tiles, int n, void *stream);
// the medium of each flat one-cell tile of a batch whose owned cells hold one
// medium in all six media arrays, else -1 (out[t] for the batch's tile t)
int launch_batch_tile_media(const BatchDev &bd, int *out, void *stream);
// the medium of each four-cell step tile whose cells (owned, halo row and
// column, planes k0..k1) hold one medium in all six media arrays, else -1
int launch_tile_media(const DevGrid &g, const StepTile *tiles, int n, int *out, void *stream);
int launch_box_gather_f32(const void *grid, size_t es, float scale, float *packed,
                          const long long ext[3], long long px, long long pxy, int xoff,
                          void *stream);
// the incident field of every TF/SF face cell at step *dn for the one-launch
// step: faces[0..nf0) act on E (time tiempo(n)), faces[nf0..nf) on H
// (tiempo(n) + dt/2); face q's cells go to inc[faces[q].inc ...], the face's
// first free axis fastest
int launch_pw_incident(const PwFace *faces, int nf0, int nf, long long maxcells,
                       const StepSrc *src, const int *dn, float *inc, void *stream);
// the pre-pass of many grids in one launch (maxnf: the most faces of a grid,
// maxcells: the most cells of a face)
int launch_pw_incident_many(const PwIncJob *jobs, int njobs, int maxnf, long long maxcells,
                            void *stream);
// increments every counter of the list (per-grid step counters of a batch)
int launch_tick_many(int *const *counters, int n, void *stream);
int launch_probes(const DevGrid &g, const ProbeDev *pr, int np, float *buf,
                  int *count, int cap, void *stream);
// the probes of several grids of a batch in one launch
struct ProbeSet {
  const ProbeDev *pr;
  float *buf;
  int *count;
  int np, cap, grid;          // grid: index into the batch's descriptors
};
int launch_probes_many(const DevGrid *grids, const ProbeSet *sets, int nsets, void *stream);
// bounding box {xmin,ymin,zmin,xmax,ymax,zmax} of the dispersive edges of
// component comp; out must hold {INT_MAX x3, INT_MIN x3} on entry
int launch_disp_bbox(const DevGrid &g, int side, int c, int32_t *out, void *stream);
int launch_mur(const DevGrid &g, const MurRun *runs, int nrun, long long total, int exact,
               void *stream);
int launch_fill_i16(int16_t *a, long long n, int16_t v, void *stream);
int launch_pec_walls(const DevGrid &g, int16_t *mi[6], int faces, void *stream);
int launch_paint(const DevGrid &g, int16_t *mi[6], int shape, const float *par,
                 int16_t medium, const int32_t *prio, int nprio, void *stream);

}  // namespace sfd
