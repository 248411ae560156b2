// fdtd_handle.h -- host-side state of one grid (the opaque fdtd_handle of
// the C ABI) and the helpers the host translation units share.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "fdtd_internal.h"
#include "sembafdtd.h"

namespace sfd_host {
using namespace sfd;
struct Source {
  int comp, hard;
  int box[6];
  float amp, deltaevol;
  int numus;
  std::vector<float> evol;
};

struct AdeIn {           // the dispersive media of one field, as set by the caller
  int ndisp = 0;
  std::vector<int32_t> dmed, dnp;
  std::vector<double> dk, db, dg3;
  long long nedges[3] = {0, 0, 0};  // dispersive edges per component
  size_t jbytes[3] = {0, 0, 0};      // pole-current bytes per component
};

struct DftIn {           // one frequency-domain observation box
  int kind;
  int box[6], stride[3];
  std::vector<float> freq;
};

// one launch of TF/SF faces: the faces of one wave and one face axis
struct PwGroup {
  int half;                       // 0: E faces, 1: H faces
  std::vector<PwRun> runs;        // step from the launch (dn = null)
  std::vector<long long> start;   // prefix of face areas, runs.size()+1 entries
  PwRun *druns[2] = {};           // device copies: [0] launch step, [1] device counter
  long long *dstart = nullptr;
};

struct PlaneWaveIn {     // sgg%PlaneWave(j) (planeonde_t, fdetypes.F90:458-465)
  int esq[6];            // esqx1, esqx2, esqy1, esqy2, esqz1, esqz2
  int nmodes;
  float p[kPwModes][3], e[kPwModes][3], incert[kPwModes];
  float deltaevol;
  std::vector<float> evol;
  float eps0, mu0;
};
}  // namespace sfd_host

struct fdtd_handle {
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;       // the stream every launch goes to
  hipStream_t solo_stream = nullptr;  // its own stream while lent to a batch
  bool own_stream = false;
  int mode = FDTD_MODE_FAST;
  int storage = FDTD_STORE_FP32;
  sfd::DevGrid g{};
  sfd::TileMap tm{};           // 64-cell x tiles (granular passes)
  sfd::TileMap tm4{};          // 256-cell x tiles (fused passes, 4 cells per lane)
  int tile_zchunk = 32;        // planes per z chunk as set (tm.kc: clipped to the grid)
  std::vector<void *> allocs;
  int64_t bytes = 0;
  float *fld[6] = {};
  int16_t *mi[6] = {};
  // media
  int nmedia = -1;
  int media_max = 0;  // largest media index written so far
  std::vector<float> g2h, gm2h;
  std::vector<int32_t> kindh;
  float *dg[4] = {};
  // spacing
  std::vector<float> idh_[6];
  float *did[6] = {};
  // cpml
  float *dprof[12] = {};
  long long psilen[12] = {};       // psi values per slab array
  // dispersive media: [0] electric (on E), [1] magnetic (on H)
  sfd_host::AdeIn adein[2];
  // first-order Mur faces
  int mur_faces[6] = {0, 0, 0, 0, 0, 0};
  std::vector<float> cab1h;            // [6][nmedia+1]
  bool mur_ready = false;
  std::vector<void *> mur_allocs;
  std::vector<sfd::MurRun> mur_runs;   // host copy (past pointers on the device)
  sfd::MurRun *dmur = nullptr;
  long long mur_total = 0;
  bool ade_ready = false;
  std::vector<void *> ade_allocs;
  // time
  double dt = 0.0;
  std::vector<double> tiempo;
  // nodal
  std::vector<sfd_host::Source> srcs;
  bool nodal_ready = false;
  std::vector<void *> nodal_allocs;
  sfd::NodalCell *dcells[2] = {};
  int ncells[2] = {0, 0};
  int serial[2] = {0, 0};
  float *dvalues = nullptr;
  int nvalues_steps = 0;
  // TF/SF plane waves
  std::vector<sfd_host::PlaneWaveIn> pwin;
  std::vector<float> line[3];           // LineX/Y/Z over [-1, n+1]
  bool pw_ready = false;
  std::vector<void *> pw_allocs;
  std::vector<sfd::PwFace> pw_faces[2];      // E faces, H faces, in reference order
  std::vector<sfd_host::PwGroup> pw_groups;  // launch groups, in reference order
  sfd::PwWave *dwaves = nullptr;
  float *dphys = nullptr;
  int pstride = 0;
  double *dtiempo = nullptr;
  // frequency-domain observation
  std::vector<sfd_host::DftIn> dftin;
  std::vector<sfd::DftDev> dfts;
  bool dft_ready = false;
  std::vector<void *> dft_allocs;     // phase tables (rebuilt with the time vector)
  std::vector<void *> dft_sums;       // running sums (kept across rebuilds)
  // probes
  sfd::ProbeDev *dprobes = nullptr;
  int nprobe = 0;
  float *dpbuf = nullptr;
  int *dpcount = nullptr;
  int pcap = 0;
  // graph
  int *dstep = nullptr;           // device step counter of the captured step
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  bool graph_ok = false;
  // the same two-pass step recorded graph_steps times in one graph
  int graph_steps = 16;
  hipGraphExec_t mexec = nullptr;
  hipGraph_t mgraph = nullptr;
  // one-launch steps (k_step): a second copy of the time-stepping state (B),
  // the A->B and B->A descriptors and a captured pair of steps
  int step_rows = 0;              // 0: two passes per step; else rows per block (4, 8, 16)
  int step_kc = 32;
  int step_creq[17] = {};  // rows per block asked for each
                                  // tile class (0: the default); a class on fewer
                                  // rows than step_rows runs the step's tiles cut
                                  // into row bands
  bool step_vec = true;           // plain tiles run four x cells per lane (k_step4)
  int lds_max = 65536;            // LDS bytes a block may use on the device
  bool step_ready = false;
  bool step_stale = true;         // B's fields must be refreshed from A
  std::vector<void *> step_allocs;       // the second state
  std::vector<void *> step_small;        // descriptors' tables
  std::vector<size_t> step_sizes;        // layout of the second state
  sfd::DevGrid gB;
  sfd::StepDev sd[2];             // [0] A -> B, [1] B -> A
  // tiles grouped by class (bit 0 CPML, 1 poles, 2 sources; 8 plain tiles
  // four cells per lane): one launch per class with tiles, each over its
  // slice of the device tile list
  std::vector<int> step_cls, step_cnt, step_off;
  std::vector<int> step_crows;           // rows per block of each class launch
  std::vector<int> step_uni;             // leading single-medium tiles of each class launch
  // per tile of the one-cell pass map tm: the medium filling it, or -1
  // (valid while tmed_ready; tiling and media changes clear it)
  int *tmed = nullptr;
  int tmed_cap = 0;
  bool tmed_ready = false;
  std::vector<double> step_cells;        // Yee cells owned per class
  std::vector<double> step_bytes;        // compulsory HBM bytes of one launch per class
  uint8_t *mi8[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};  // step mode
  void *stage = nullptr;                 // packed staging buffer of host transfers
  size_t stage_bytes = 0;
  // the stored previous field of the pole updates may differ from the field
  // (fields set from outside, granular calls): the next pair reads it
  bool prev_exact = true;
  bool prev_lazy = false;      // the stored previous fields await settle_prev
  sfd::StepTile *dtiles = nullptr;
  std::vector<sfd::StepTile> step_tiles;  // host copy of the tile list (batches)
  // TF/SF faces of both halves for the step kernels and the buffer of their
  // incident field, refreshed before every step
  sfd::PwFace *dsfaces = nullptr;
  int step_nf0 = 0, step_nf = 0;
  long long step_maxcells = 0;
  float *dinc = nullptr;
  sfd::StepSrc *dstepsrc = nullptr;
  hipGraphExec_t pexec = nullptr;
  hipGraph_t pgraph = nullptr;
  std::vector<sfd::NodalCell> hcells[2];   // host copies of the nodal cells
  unsigned version = 0;           // bumped by every change that drops the graph
  // the batch the grid is admitted to (null: none).  While it is set the
  // batch's captured step replays over the grid's buffers and shares its step
  // counter, so the grid's own stepping calls and fdtd_destroy are refused.
  fdtd_batch *batch = nullptr;
  // probe samples taken since the last fdtd_read_probes (or fdtd_set_probes):
  // a call that would take more than the buffer holds is refused
  long long pbuffered = 0;
};


// helpers shared by the host translation units (fdtd_host.cpp)
namespace sfd_host {
int fail(fdtd_handle *h, int code, const char *fmt, ...) __attribute__((format(printf, 3, 4)));
int dalloc(fdtd_handle *h, void **p, size_t bytes, std::vector<void *> *list = nullptr);
int refuse_batched(fdtd_handle *h);
// extents of component comp's Alloc box per axis
void alloc_extent(const fdtd_handle *h, int comp, long long ext[3]);
int probe_room(fdtd_handle *h, long long more);
int prepare(fdtd_handle *h);
int build_tile_media(fdtd_handle *h);
int run_pw(fdtd_handle *h, int half, int n, const int *dn);
int run_nodal(fdtd_handle *h, int half, int n, const int *dn);
int run_dft(fdtd_handle *h, int n, const int *dn);
int run_mur(fdtd_handle *h);
int run_ade(fdtd_handle *h, int side);
void invalidate_graph(fdtd_handle *h);
// the one-launch step: its descriptors and tile plan (step_rows > 0), the
// second state refreshed from the first when stale, a class's rows per block
int build_step(fdtd_handle *h);
int sync_b(fdtd_handle *h);
int class_rows(const fdtd_handle *h, int c);
bool step_supported(const fdtd_handle *h, int rows);
}  // namespace sfd_host

// a failed HIP call on grid h: its message, and FDTD_ERR_HIP from the caller
#define HIPCHK(h, call)                                                        \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      return sfd_host::fail(h, FDTD_ERR_HIP, "%s: %s (%s:%d)", #call,          \
                            hipGetErrorString(e_), __FILE__, __LINE__);        \
  } while (0)
