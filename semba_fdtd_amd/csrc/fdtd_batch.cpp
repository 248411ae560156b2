// fdtd_batch.cpp -- batches of independent grids advanced together.
//
// A batch holds grids that were set up as single grids (fdtd_create and the
// set-up calls); admitting one hands its device state to the batch without
// copying.  One step of the batch runs, for every grid at its own step:
//   E face pass of all grids          one persistent launch (k_pass_batch)
//   pole stage                        one launch for the grids with poles
//   plane-wave E faces                one launch per (wave, face axis) group
//   nodal E                           one launch per grid with E sources
//   H face pass of all grids          one persistent launch
//   plane-wave H faces, nodal H       as above
//   probes                            one launch for all grids
//   frequency-domain sums             per grid
//   step counters                     one launch for all grids
// captured once per membership in a HIP graph and replayed.  Each grid keeps
// its own step counter (its dstep), so grids admitted at different times
// advance in lock-step from different n.
//
// With one launch per step (fdtd_batch_set_step_mode) the batch advances
// pairs of steps (A -> B -> A, every grid holding its second state copy), and
// each step of a pair runs:
//   incident field of the TF/SF faces  one launch for all grids with faces
//   step kernels                       one launch per tile class (and block
//                                      shape and media width) over the tiles
//                                      of every grid's own tile plan
//   Mur faces                          per grid with Mur faces
//   probes, step counters              one launch each for all grids
//   frequency-domain sums              per grid
// An odd last step runs the two-pass step above.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "fdtd_handle.h"

using namespace sfd;
using namespace sfd_host;

namespace {
thread_local std::string g_batch_error;

// the plane-wave groups of one (half, wave, axis) key across the batch
struct BatchPwGroup {
  int half;
  std::vector<PwRun> runs;
  std::vector<long long> start;
  PwRun *druns = nullptr;
  long long *dstart = nullptr;
};
}  // namespace

struct fdtd_batch {
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int mode = -1;
  int fp16 = 0;                       // field storage shared by the grids
  std::vector<fdtd_handle *> slots;   // admitted grids; retired slots are null
  std::vector<long long> step;        // host copy of each slot's step counter
  std::vector<unsigned> version;      // each grid's state version at the last build
  int rows = 4, zchunk = 16, xcells = 1, blocks = 0;
  // device descriptors, rebuilt when the membership changes
  bool ready = false;
  std::vector<void *> allocs;
  BatchDev bd{};
  size_t lds = 0;
  std::vector<BatchPwGroup> pw;
  int **dcounters = nullptr;
  int ncounters = 0;
  // grids with pole state per field (indices into bd.grids) for the one
  // pole-stage launch per half, and the probe sets of all grids
  int *dade[2] = {nullptr, nullptr};
  int nade[2] = {0, 0};
  size_t ade_lds[2] = {0, 0};
  int ade_tbl[2] = {1, 1};
  long long ade_nmax[2] = {0, 0};
  ProbeSet *dprobes = nullptr;
  int nprobesets = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  // one launch per step: rows per block (0: two passes) and planes per chunk
  int step_rows = 0, step_kc = 16;
  bool sready = false;
  std::vector<void *> sallocs;
  StepDev *dsd[2][2] = {};           // [direction][previous field exact]: per grid
  DevGrid *dgo[2] = {};              // per direction: the written copy of each grid
  struct ClassLaunch {
    int cls, narrow, rows;
    size_t lds;
    GridTile *tiles;
    int n;
  };
  std::vector<ClassLaunch> slaunch;
  PwIncJob *djobs = nullptr;
  int njobs = 0, maxnf = 0;
  long long maxcells = 0;
  hipGraph_t pgraph = nullptr;
  hipGraphExec_t pexec = nullptr;
};

namespace {

int bfail(fdtd_batch *b, int code, const char *fmt, ...) __attribute__((format(printf, 3, 4)));
int bfail(fdtd_batch *b, int code, const char *fmt, ...) {
  (void)hipGetLastError();
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (b) b->err = buf; else g_batch_error = buf;
  return code;
}

#define BCHK(b, call)                                                          \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      return bfail(b, FDTD_ERR_HIP, "%s: %s (%s:%d)", #call,                   \
                   hipGetErrorString(e_), __FILE__, __LINE__);                 \
  } while (0)

void drop_graph(fdtd_batch *b) {
  if (b->gexec) hipGraphExecDestroy(b->gexec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
}

void drop_step(fdtd_batch *b) {
  if (b->pexec) hipGraphExecDestroy(b->pexec);
  if (b->pgraph) hipGraphDestroy(b->pgraph);
  b->pexec = nullptr;
  b->pgraph = nullptr;
  for (void *p : b->sallocs) hipFree(p);
  b->sallocs.clear();
  b->slaunch.clear();
  b->sready = false;
}

void drop_descriptors(fdtd_batch *b) {
  drop_graph(b);
  drop_step(b);
  for (void *p : b->allocs) hipFree(p);
  b->allocs.clear();
  b->pw.clear();
  b->ready = false;
}

template <class T>
int bupload(fdtd_batch *b, T **dst, const T *src, size_t n, std::vector<void *> *list = nullptr) {
  void *p = nullptr;
  BCHK(b, hipMalloc(&p, std::max<size_t>(n * sizeof(T), 16)));
  (list ? list : &b->allocs)->push_back(p);
  if (n) BCHK(b, hipMemcpyAsync(p, src, n * sizeof(T), hipMemcpyHostToDevice, b->stream));
  *dst = (T *)p;
  return FDTD_OK;
}

// tile map of one grid under the batch's tiling: four x cells per lane in
// 256-wide segments, or one cell per lane in flat tiles (runs of 64 * rows
// columns of the x-y plane, x fastest), which leave no lane idle on grids a
// few tens of cells wide
TileMap batch_tiles(const fdtd_batch *b, const DevGrid &g) {
  TileMap tm{};
  tm.bdy = b->rows;
  if (b->xcells == 4) {
    const int w = 4 * kWave;
    tm.vec = 4;
    tm.tx = (g.nx + 1 + w - 1) / w;
    tm.ty = (g.ny + 1 + b->rows - 1) / b->rows;
  } else {
    const long long cols = (long long)(g.nx + 1) * (g.ny + 1), per = (long long)kWave * b->rows;
    tm.vec = 1;
    tm.flat = 1;
    tm.tx = (int)((cols + per - 1) / per);
    tm.ty = 1;
  }
  tm.kc = std::min(b->zchunk, g.nz + 1);
  tm.tz = (g.nz + 1 + tm.kc - 1) / tm.kc;
  tm.nblk = tm.tx * tm.ty * tm.tz;
  tm.remap = 0;
  return tm;
}

int build_descriptors(fdtd_batch *b) {
  if (b->ready) return FDTD_OK;
  drop_descriptors(b);
  std::vector<DevGrid> grids;
  std::vector<TileMap> tms;
  std::vector<int> start(1, 0);
  std::vector<int *> counters;
  size_t lds = 16;
  int maxwaves = 0;
  int rc;
  for (size_t s = 0; s < b->slots.size(); ++s) {
    fdtd_handle *h = b->slots[s];
    if (!h) continue;
    // a grid changed since admission (set-up calls) rebuilds its device state
    if ((rc = prepare(h))) return bfail(b, rc, "slot %zu: %s", s, h->err.c_str());
    b->version[s] = h->version;
    grids.push_back(h->g);
    tms.push_back(batch_tiles(b, h->g));
    if ((long long)start.back() + tms.back().nblk > 0x7fffffffLL)
      return bfail(b, FDTD_ERR_ARG, "batch holds too many tiles");
    start.push_back(start.back() + tms.back().nblk);
    counters.push_back(h->dstep);
    lds = std::max(lds, (size_t)(h->g.nmedia + 1) * 12);
    maxwaves = std::max(maxwaves, (int)h->pwin.size());
  }
  DevGrid *dg = nullptr;
  TileMap *dt = nullptr;
  int *ds = nullptr;
  if ((rc = bupload(b, &dg, grids.data(), grids.size())) ||
      (rc = bupload(b, &dt, tms.data(), tms.size())) ||
      (rc = bupload(b, &ds, start.data(), start.size())) ||
      (rc = bupload(b, &b->dcounters, counters.data(), counters.size())))
    return rc;
  b->bd.grids = dg;
  b->bd.tms = dt;
  b->bd.start = ds;
  b->bd.ngrid = (int)grids.size();
  b->bd.total = start.back();
  // flat tiles whose cells all hold one medium skip the media loads
  b->bd.tmed = nullptr;
  if (b->xcells == 1 && b->bd.total > 0) {
    int *dm = nullptr;
    BCHK(b, hipMalloc((void **)&dm, (size_t)b->bd.total * sizeof(int)));
    b->allocs.push_back(dm);
    BCHK(b, (hipError_t)launch_batch_tile_media(b->bd, dm, b->stream));
    b->bd.tmed = dm;
  }
  b->ncounters = (int)counters.size();
  b->lds = lds;
  // pole stage and probes: one launch each for the whole batch
  {
    const bool exact = b->mode == FDTD_MODE_EXACT;
    std::vector<ProbeSet> sets;
    std::vector<int> al[2];
    for (int side = 0; side < 2; ++side) {
      b->ade_lds[side] = 0;
      b->ade_tbl[side] = 1;
      b->ade_nmax[side] = 0;
    }
    int gi = 0;
    for (size_t s = 0; s < b->slots.size(); ++s) {
      fdtd_handle *h = b->slots[s];
      if (!h) continue;
      for (int side = 0; side < 2; ++side) {
        const AdeSet &ad = h->g.ade[side];
        if (!ad.has) continue;
        long long nmax = 0;
        for (int c = 0; c < 3; ++c) nmax = std::max(nmax, ad.nstate[c]);
        if (nmax <= 0) continue;
        al[side].push_back(gi);
        b->ade_nmax[side] = std::max(b->ade_nmax[side], nmax);
        const size_t la = exact ? exact::ade_lds_of(h->g, side) : fast::ade_lds_of(h->g, side);
        b->ade_lds[side] = std::max(b->ade_lds[side], la);
        if (h->g.nmedia + 1 > 4096) b->ade_tbl[side] = 0;
      }
      if (h->nprobe) {
        ProbeSet ps{};
        ps.pr = h->dprobes;
        ps.buf = h->dpbuf;
        ps.count = h->dpcount;
        ps.np = h->nprobe;
        ps.cap = h->pcap;
        ps.grid = gi;
        sets.push_back(ps);
      }
      ++gi;
    }
    for (int side = 0; side < 2; ++side) {
      if (b->ade_lds[side] > 48 * 1024) b->ade_tbl[side] = 0;
      b->nade[side] = (int)al[side].size();
      b->dade[side] = nullptr;
      if (!al[side].empty() && (rc = bupload(b, &b->dade[side], al[side].data(), al[side].size())))
        return rc;
    }
    b->nprobesets = (int)sets.size();
    b->dprobes = nullptr;
    if (!sets.empty() && (rc = bupload(b, &b->dprobes, sets.data(), sets.size()))) return rc;
  }
  // plane-wave groups across the batch, in each grid's face order: per half,
  // per wave index, per face axis (the runs read each grid's own counter)
  for (int half = 0; half < 2; ++half)
    for (int w = 0; w < maxwaves; ++w)
      for (int axis = 0; axis < 3; ++axis) {
        BatchPwGroup grp;
        grp.half = half;
        grp.start.push_back(0);
        for (fdtd_handle *h : b->slots) {
          if (!h) continue;
          for (const PwGroup &g : h->pw_groups) {
            if (g.half != half || g.runs.empty() || g.runs[0].f.wave != w ||
                g.runs[0].f.axis != axis)
              continue;
            for (size_t r = 0; r < g.runs.size(); ++r) {
              PwRun run = g.runs[r];
              run.dn = h->dstep;
              grp.runs.push_back(run);
              grp.start.push_back(grp.start.back() + (g.start[r + 1] - g.start[r]));
            }
          }
        }
        if (grp.runs.empty()) continue;
        if ((rc = bupload(b, &grp.druns, grp.runs.data(), grp.runs.size())) ||
            (rc = bupload(b, &grp.dstart, grp.start.data(), grp.start.size())))
          return rc;
        b->pw.push_back(std::move(grp));
      }
  BCHK(b, hipStreamSynchronize(b->stream));
  b->ready = true;
  return FDTD_OK;
}

// The one-launch step of every grid: each grid's own tile plan and
// descriptors (build_step, under the batch's rows and chunk), gathered into
// one launch per (class, block rows, media width).
int build_step_descriptors(fdtd_batch *b) {
  if (b->sready) return FDTD_OK;
  drop_step(b);
  const bool exact = b->mode == FDTD_MODE_EXACT;
  int rc;
  std::vector<StepDev> sds[2][2];
  std::vector<DevGrid> go[2];
  std::vector<PwIncJob> jobs;
  struct Key {
    int cls, narrow, rows;
    bool operator<(const Key &o) const {
      return cls != o.cls ? cls < o.cls : narrow != o.narrow ? narrow < o.narrow : rows < o.rows;
    }
  };
  std::vector<std::pair<Key, std::vector<GridTile>>> groups;
  std::vector<size_t> glds;
  b->maxnf = 0;
  b->maxcells = 0;
  int gi = 0;
  for (size_t s = 0; s < b->slots.size(); ++s) {
    fdtd_handle *h = b->slots[s];
    if (!h) continue;
    if (!step_supported(h, b->step_rows))
      return bfail(b, FDTD_ERR_STATE,
                   "slot %zu: one-launch steps need fp32 storage, small media tables and at "
                   "most 64 TF/SF faces per half", s);
    if (h->step_rows != b->step_rows || h->step_kc != b->step_kc) {
      h->step_rows = b->step_rows;
      h->step_kc = b->step_kc;
      invalidate_graph(h);
    }
    if ((rc = build_step(h)) || (rc = sync_b(h))) return bfail(b, rc, "slot %zu: %s", s, h->err.c_str());
    for (int d = 0; d < 2; ++d) {
      for (int e = 0; e < 2; ++e) {
        StepDev sd = h->sd[d];
        sd.prev_from_field = e ? 0 : 1;
        sd.tiles = nullptr;
        sd.tm.nblk = 0;
        sds[d][e].push_back(sd);
      }
      go[d].push_back(h->sd[d].go);
    }
    if (h->step_nf > 0) {
      jobs.push_back(PwIncJob{h->dsfaces, h->step_nf0, h->step_nf, h->dstepsrc, h->dstep, h->dinc});
      b->maxnf = std::max(b->maxnf, h->step_nf);
      b->maxcells = std::max(b->maxcells, h->step_maxcells);
    }
    const DevGrid &g = h->g;
    for (size_t q = 0; q < h->step_cls.size(); ++q) {
      const Key key{h->step_cls[q], h->sd[0].narrow, h->step_crows[q]};
      size_t gq = 0;
      while (gq < groups.size() && (groups[gq].first < key || key < groups[gq].first)) ++gq;
      if (gq == groups.size()) {
        groups.push_back({key, {}});
        glds.push_back(0);
      }
      const size_t lds =
          key.cls == 8 ? (exact ? exact::step4_lds_bytes(g, key.rows) : fast::step4_lds_bytes(g, key.rows))
                       : (exact ? exact::step_lds_bytes(g, key.rows, h->step_nf, h->sd[0].tm.kc)
                                : fast::step_lds_bytes(g, key.rows, h->step_nf, h->sd[0].tm.kc));
      glds[gq] = std::max(glds[gq], lds);
      for (int t = 0; t < h->step_cnt[q]; ++t) {
        GridTile gt{};
        gt.t = h->step_tiles[(size_t)h->step_off[q] + (size_t)t];
        gt.grid = gi;
        groups[gq].second.push_back(gt);
      }
    }
    ++gi;
  }
  for (int d = 0; d < 2; ++d) {
    for (int e = 0; e < 2; ++e)
      if ((rc = bupload(b, &b->dsd[d][e], sds[d][e].data(), sds[d][e].size(), &b->sallocs)))
        return rc;
    if ((rc = bupload(b, &b->dgo[d], go[d].data(), go[d].size(), &b->sallocs))) return rc;
  }
  b->njobs = (int)jobs.size();
  b->djobs = nullptr;
  if (!jobs.empty() && (rc = bupload(b, &b->djobs, jobs.data(), jobs.size(), &b->sallocs)))
    return rc;
  for (size_t gq = 0; gq < groups.size(); ++gq) {
    const std::vector<GridTile> &v = groups[gq].second;
    if (v.empty()) continue;
    if (v.size() > 0x7fffffffULL) return bfail(b, FDTD_ERR_ARG, "batch holds too many tiles");
    fdtd_batch::ClassLaunch cl{groups[gq].first.cls, groups[gq].first.narrow,
                               groups[gq].first.rows, glds[gq], nullptr, (int)v.size()};
    if ((rc = bupload(b, &cl.tiles, v.data(), v.size(), &b->sallocs))) return rc;
    b->slaunch.push_back(cl);
  }
  BCHK(b, hipStreamSynchronize(b->stream));
  b->sready = true;
  return FDTD_OK;
}

// one step of a pair for every grid, direction d (0: A -> B, 1: B -> A)
int launch_step_dir(fdtd_batch *b, int d, bool exact_prev) {
  const bool exact = b->mode == FDTD_MODE_EXACT;
  int e, rc;
  if (b->njobs &&
      launch_pw_incident_many(b->djobs, b->njobs, b->maxnf, b->maxcells, b->stream))
    return bfail(b, FDTD_ERR_HIP, "batch incident-field launch failed");
  const StepDev *sds = b->dsd[d][exact_prev ? 1 : 0];
  for (const fdtd_batch::ClassLaunch &cl : b->slaunch) {
    e = exact ? exact::launch_step_many(sds, cl.tiles, cl.n, cl.narrow, cl.rows, cl.cls, cl.lds,
                                        b->stream)
              : fast::launch_step_many(sds, cl.tiles, cl.n, cl.narrow, cl.rows, cl.cls, cl.lds,
                                       b->stream);
    if (e) return bfail(b, FDTD_ERR_HIP, "batch step launch failed: %s", hipGetErrorString((hipError_t)e));
  }
  for (fdtd_handle *h : b->slots) {
    if (!h || h->mur_runs.empty()) continue;
    if (launch_mur(h->sd[d].go, h->dmur, (int)h->mur_runs.size(), h->mur_total, exact, b->stream))
      return bfail(b, FDTD_ERR_HIP, "batch Mur launch failed");
  }
  if (launch_probes_many(b->dgo[d], b->dprobes, b->nprobesets, b->stream))
    return bfail(b, FDTD_ERR_HIP, "batch probe launch failed");
  for (fdtd_handle *h : b->slots) {
    if (!h) continue;
    for (const DftDev &df : h->dfts)
      if (launch_dft(h->sd[d].go, df, 0, h->dstep, b->stream))
        return bfail(b, FDTD_ERR_HIP, "batch DFT launch failed");
  }
  (void)rc;
  if (launch_tick_many(b->dcounters, b->ncounters, b->stream))
    return bfail(b, FDTD_ERR_HIP, "batch step counter launch failed");
  return FDTD_OK;
}

int launch_step(fdtd_batch *b) {
  const bool exact = b->mode == FDTD_MODE_EXACT;
  // persistent blocks: 16384 measured fastest on the 256 uneven instances
  // (flat tiles of 64 x 4 threads, 16-plane chunks)
  const int blocks = b->blocks > 0 ? b->blocks : 16384;
  int e, rc;
  for (int half = 0; half < 2; ++half) {
    const int vec = b->xcells == 4 ? 4 : 1;
    e = exact ? exact::launch_pass_batch(b->bd, half, blocks, b->rows, vec, b->lds, 0, b->stream)
              : fast::launch_pass_batch(b->bd, half, blocks, b->rows, vec, b->lds, b->fp16,
                                        b->stream);
    if (e) return bfail(b, FDTD_ERR_HIP, "batch face pass failed: %s", hipGetErrorString((hipError_t)e));
    // the pole stage of every grid with poles on this field follows the face
    // pass, in one launch
    if (b->nade[half]) {
      e = exact ? exact::launch_ade_many(b->bd.grids, b->dade[half], b->nade[half], half,
                                         b->ade_tbl[half] ? b->ade_lds[half] : 0,
                                         b->ade_tbl[half], b->ade_nmax[half], b->stream)
                : fast::launch_ade_many(b->bd.grids, b->dade[half], b->nade[half], half,
                                        b->ade_tbl[half] ? b->ade_lds[half] : 0,
                                        b->ade_tbl[half], b->ade_nmax[half], b->stream);
      if (e) return bfail(b, FDTD_ERR_HIP, "batch pole stage failed: %s", hipGetErrorString((hipError_t)e));
    }
    for (const BatchPwGroup &g : b->pw) {
      if (g.half != half) continue;
      e = launch_tfsf_list(g.druns, g.dstart, (int)g.runs.size(), g.start.back(), 0, b->stream);
      if (e) return bfail(b, FDTD_ERR_HIP, "batch plane-wave launch failed");
    }
    for (fdtd_handle *h : b->slots)
      if (h && (rc = run_nodal(h, half, 0, h->dstep)))
        return bfail(b, rc, "grid: %s", h->err.c_str());
  }
  for (fdtd_handle *h : b->slots)
    if (h && (rc = run_mur(h))) return bfail(b, rc, "grid: %s", h->err.c_str());
  if (launch_probes_many(b->bd.grids, b->dprobes, b->nprobesets, b->stream))
    return bfail(b, FDTD_ERR_HIP, "batch probe launch failed");
  for (fdtd_handle *h : b->slots)
    if (h && (rc = run_dft(h, 0, h->dstep))) return bfail(b, rc, "grid: %s", h->err.c_str());
  if (launch_tick_many(b->dcounters, b->ncounters, b->stream))
    return bfail(b, FDTD_ERR_HIP, "batch step counter launch failed");
  return FDTD_OK;
}

}  // namespace

extern "C" {

const char *fdtd_batch_last_error(const fdtd_batch *b) {
  return b ? b->err.c_str() : g_batch_error.c_str();
}

int fdtd_batch_create(int32_t device, void *stream, fdtd_batch **out) {
  if (!out) return bfail(nullptr, FDTD_ERR_ARG, "null argument");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return bfail(nullptr, FDTD_ERR_DEVICE, "no HIP device available");
  if (device < 0 || device >= ndev)
    return bfail(nullptr, FDTD_ERR_DEVICE, "device %d out of range", device);
  fdtd_batch *b = new fdtd_batch();
  b->device = device;
  hipSetDevice(device);
  if (stream) {
    b->stream = (hipStream_t)stream;
  } else {
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
      delete b;
      return bfail(nullptr, FDTD_ERR_HIP, "hipStreamCreate failed");
    }
    b->own_stream = true;
  }
  *out = b;
  return FDTD_OK;
}

int fdtd_batch_destroy(fdtd_batch *b) {
  if (!b) return FDTD_OK;
  hipSetDevice(b->device);
  hipStreamSynchronize(b->stream);
  for (fdtd_handle *h : b->slots)
    if (h) {
      if (h->solo_stream) {
        h->stream = h->solo_stream;
        h->solo_stream = nullptr;
      }
      h->batch = nullptr;
      invalidate_graph(h);
    }
  drop_descriptors(b);
  if (b->own_stream) hipStreamDestroy(b->stream);
  delete b;
  return FDTD_OK;
}

int fdtd_batch_admit(fdtd_batch *b, fdtd_handle *h, int32_t n0, int32_t *slot) {
  if (!b || !h || !slot || n0 < 0) return bfail(b, FDTD_ERR_ARG, "bad admit arguments");
  if (h->device != b->device) return bfail(b, FDTD_ERR_ARG, "grid lives on another device");
  // a grid belongs to one batch at a time (this one or another)
  if (h->batch) return bfail(b, FDTD_ERR_STATE, "grid already admitted to a batch");
  if (b->mode >= 0 && (h->mode != b->mode || h->g.fp16 != b->fp16))
    return bfail(b, FDTD_ERR_ARG, "grids of one batch share the arithmetic mode and storage");
  hipSetDevice(b->device);
  int rc = prepare(h);
  if (rc) return bfail(b, rc, "grid not ready: %s", h->err.c_str());
  if (h->nmedia + 1 > 4096) return bfail(b, FDTD_ERR_ARG, "batched grids hold at most 4095 media");
  // the grid's work is ordered after anything still queued on its own stream
  BCHK(b, hipStreamSynchronize(h->stream));
  BCHK(b, hipMemcpyAsync(h->dstep, &n0, 4, hipMemcpyHostToDevice, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  invalidate_graph(h);
  // while admitted, the grid's own launches (set-up, reads) go to the batch's
  // stream, so they are ordered with the batch steps
  h->solo_stream = h->stream;
  h->stream = b->stream;
  b->mode = h->mode;
  b->fp16 = h->g.fp16;
  int s = -1;
  for (size_t i = 0; i < b->slots.size(); ++i)
    if (!b->slots[i]) { s = (int)i; break; }
  if (s < 0) {
    s = (int)b->slots.size();
    b->slots.push_back(nullptr);
    b->step.push_back(0);
    b->version.push_back(0);
  }
  b->slots[s] = h;
  h->batch = b;
  b->step[s] = n0;
  b->ready = false;
  *slot = s;
  return FDTD_OK;
}

int fdtd_batch_retire(fdtd_batch *b, int32_t slot) {
  if (!b || slot < 0 || slot >= (int32_t)b->slots.size() || !b->slots[slot])
    return bfail(b, FDTD_ERR_ARG, "no grid in slot %d", slot);
  hipSetDevice(b->device);
  BCHK(b, hipStreamSynchronize(b->stream));
  fdtd_handle *h = b->slots[slot];
  if (h->solo_stream) {
    h->stream = h->solo_stream;
    h->solo_stream = nullptr;
  }
  invalidate_graph(h);
  h->batch = nullptr;
  b->slots[slot] = nullptr;
  b->ready = false;
  bool any = false;
  for (fdtd_handle *h : b->slots) any = any || h;
  if (!any) b->mode = -1;
  return FDTD_OK;
}

int64_t fdtd_batch_step_of(const fdtd_batch *b, int32_t slot) {
  if (!b || slot < 0 || slot >= (int32_t)b->slots.size() || !b->slots[slot]) return -1;
  return b->step[slot];
}

int fdtd_batch_set_tiling(fdtd_batch *b, int32_t rows, int32_t zchunk, int32_t xcells,
                          int32_t blocks) {
  if (!b || !(rows == 1 || rows == 2 || rows == 4) || zchunk < 1 ||
      !(xcells == 1 || xcells == 4) || blocks < 0)
    return bfail(b, FDTD_ERR_ARG, "tiling %d rows x %d planes x %d cells, %d blocks not supported",
                 rows, zchunk, xcells, blocks);
  b->rows = rows;
  b->zchunk = zchunk;
  b->xcells = xcells;
  b->blocks = blocks;
  b->ready = false;
  return FDTD_OK;
}

int fdtd_batch_set_step_mode(fdtd_batch *b, int32_t rows, int32_t zchunk) {
  if (!b || !(rows == 0 || rows == 4 || rows == 8 || rows == 16) || zchunk < 1)
    return bfail(b, FDTD_ERR_ARG, "batch step mode %d rows x %d planes not supported", rows,
                 zchunk);
  b->step_rows = rows;
  b->step_kc = zchunk;
  drop_step(b);
  return FDTD_OK;
}

int fdtd_batch_step(fdtd_batch *b, int32_t nsteps) {
  if (!b || nsteps < 0) return bfail(b, FDTD_ERR_ARG, "bad step count");
  hipSetDevice(b->device);
  bool any = false;
  for (size_t s = 0; s < b->slots.size(); ++s) {
    fdtd_handle *h = b->slots[s];
    if (!h) continue;
    any = true;
    const long long last = b->step[s] + nsteps;
    if ((!h->srcs.empty() && last > h->nvalues_steps) ||
        ((!h->pwin.empty() || !h->dftin.empty()) && last > (long long)h->tiempo.size()))
      return bfail(b, FDTD_ERR_ARG, "slot %zu: steps up to %lld beyond its time vector", s,
                   last - 1);
    if (h->nprobe && h->pbuffered + nsteps > h->pcap)
      return bfail(b, FDTD_ERR_ARG,
                   "slot %zu: probe buffer holds %d steps, %lld buffered, %d more asked", s,
                   h->pcap, h->pbuffered, nsteps);
  }
  if (!any || nsteps == 0) return FDTD_OK;
  for (size_t s = 0; s < b->slots.size(); ++s)
    if (b->slots[s] && b->slots[s]->version != b->version[s]) b->ready = false;
  int rc = build_descriptors(b);
  if (rc) return rc;
  // one launch per step: pairs A -> B -> A, an odd last step on the two passes
  const int pairs = b->step_rows > 0 ? nsteps / 2 : 0;
  const int singles = nsteps - 2 * pairs;
  if (pairs > 0) {
    if ((rc = build_step_descriptors(b))) return rc;
    // (building the members' step state bumped their versions)
    for (size_t s = 0; s < b->slots.size(); ++s)
      if (b->slots[s]) b->version[s] = b->slots[s]->version;
    bool exact_prev = false;
    for (fdtd_handle *h : b->slots)
      if (h) {
        // fields written from outside since the last pair: B follows A
        if ((rc = sync_b(h))) return bfail(b, rc, "grid: %s", h->err.c_str());
        exact_prev = exact_prev || h->prev_exact;
      }
    int p0 = 0;
    if (exact_prev) {
      // the first pair after fields were set from outside reads the stored
      // previous fields of the pole updates
      if ((rc = launch_step_dir(b, 0, true)) || (rc = launch_step_dir(b, 1, true))) return rc;
      p0 = 1;
      for (fdtd_handle *h : b->slots)
        if (h) h->prev_exact = false;
    }
    if (p0 < pairs && !b->pexec) {
      BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      int r2 = launch_step_dir(b, 0, false);
      if (!r2) r2 = launch_step_dir(b, 1, false);
      hipGraph_t graph = nullptr;
      hipError_t ec = hipStreamEndCapture(b->stream, &graph);
      if (r2) { if (graph) hipGraphDestroy(graph); return r2; }
      if (ec != hipSuccess) return bfail(b, FDTD_ERR_HIP, "batch capture failed: %s", hipGetErrorString(ec));
      b->pgraph = graph;
      BCHK(b, hipGraphInstantiate(&b->pexec, graph, nullptr, nullptr, 0));
    }
    for (int q = p0; q < pairs; ++q) BCHK(b, hipGraphLaunch(b->pexec, b->stream));
  }
  if (singles > 0 && !b->gexec) {
    BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int r2 = launch_step(b);
    hipGraph_t graph = nullptr;
    hipError_t ec = hipStreamEndCapture(b->stream, &graph);
    if (r2) { if (graph) hipGraphDestroy(graph); return r2; }
    if (ec != hipSuccess) return bfail(b, FDTD_ERR_HIP, "batch capture failed: %s", hipGetErrorString(ec));
    b->graph = graph;
    BCHK(b, hipGraphInstantiate(&b->gexec, graph, nullptr, nullptr, 0));
  }
  for (int32_t s = 0; s < singles; ++s) BCHK(b, hipGraphLaunch(b->gexec, b->stream));
  for (size_t s = 0; s < b->slots.size(); ++s)
    if (b->slots[s]) {
      b->step[s] += nsteps;
      b->slots[s]->pbuffered += nsteps;
    }
  return FDTD_OK;
}

int fdtd_batch_synchronize(fdtd_batch *b) {
  if (!b) return FDTD_ERR_ARG;
  hipSetDevice(b->device);
  BCHK(b, hipStreamSynchronize(b->stream));
  return FDTD_OK;
}

}  // extern "C"
