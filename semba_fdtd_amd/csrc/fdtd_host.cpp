// fdtd_host.cpp -- C ABI of the gfx950 field time-stepping library: device
// memory, set-up uploads, launches and the captured step graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "fdtd_handle.h"
#include "fdtd_plan.h"

using namespace sfd;
using namespace sfd_host;

namespace {
thread_local std::string g_create_error;
}  // namespace

namespace sfd_host {

int fail(fdtd_handle *h, int code, const char *fmt, ...) {
  (void)hipGetLastError();  // a failed call must not surface again in the next launch check
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (h) h->err = buf; else g_create_error = buf;
  return code;
}


int dalloc(fdtd_handle *h, void **p, size_t bytes, std::vector<void *> *list) {
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess)
    return fail(h, FDTD_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  (list ? list : &h->allocs)->push_back(*p);
  h->bytes += (int64_t)bytes;
  return FDTD_OK;
}

int dzero(fdtd_handle *h, void *p, size_t bytes) {
  HIPCHK(h, hipMemsetAsync(p, 0, bytes, h->stream));
  return FDTD_OK;
}

template <class T>
int upload(fdtd_handle *h, T **dst, const T *src, size_t n,
           std::vector<void *> *list = nullptr) {
  if (!*dst) {
    int rc = dalloc(h, (void **)dst, n * sizeof(T), list);
    if (rc) return rc;
  }
  HIPCHK(h, hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FDTD_OK;
}

size_t field_elems(const fdtd_handle *h) { return (size_t)h->g.pxy * (size_t)(h->g.nz + 3); }

// bytes per stored field value
size_t field_bytes(const fdtd_handle *h) { return h->storage == FDTD_STORE_FP16 ? 2 : 4; }

int axis_n(const fdtd_handle *h, int a) { return a == 0 ? h->g.nx : (a == 1 ? h->g.ny : h->g.nz); }

// extents of component comp's Alloc box [-1, last swept + 1] per axis, the
// shape launch_simulation allocates (timestepping.F90:380-386)
void alloc_extent(const fdtd_handle *h, int comp, long long ext[3]) {
  for (int a = 0; a < 3; ++a) {
    const bool node = comp < 3 ? (comp != a) : (comp - 3 == a);
    ext[a] = (long long)axis_n(h, a) + (node ? 3 : 2);
  }
}

// host Alloc-box array <-> device grid array (element size es), one 2-D copy
// per z plane
// A component's Alloc box between the host (dense, x fastest) and its grid
// array: contiguous copies of runs of z planes through a packed device
// staging buffer (at most 64 MiB, or one plane), and a kernel that moves each
// run into (or out of) the padded layout.
int copy_box(fdtd_handle *h, int comp, void *dev, const void *host_src, void *host_dst,
             size_t es) {
  long long ext[3];
  alloc_extent(h, comp, ext);
  const size_t plane = (size_t)(ext[0] * ext[1]) * es;
  const long long per = std::max<long long>(1, (long long)((64u << 20) / std::max<size_t>(plane, 1)));
  const size_t want = plane * (size_t)std::min(per, ext[2]);
  if (want > h->stage_bytes) {
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (h->stage) hipFree(h->stage);
    h->stage = nullptr;
    h->stage_bytes = 0;
    HIPCHK(h, hipMalloc(&h->stage, want));
    h->stage_bytes = want;
  }
  for (long long k0 = 0; k0 < ext[2]; k0 += per) {
    const long long nk = std::min(per, ext[2] - k0);
    const long long sub[3] = {ext[0], ext[1], nk};
    char *g = (char *)dev + (size_t)k0 * (size_t)h->g.pxy * es;
    const size_t bytes = plane * (size_t)nk;
    if (host_src) {
      HIPCHK(h, hipMemcpyAsync(h->stage, (const char *)host_src + (size_t)k0 * plane, bytes,
                               hipMemcpyHostToDevice, h->stream));
      HIPCHK(h, (hipError_t)launch_box_move(g, h->stage, es, sub, h->g.px, h->g.pxy, h->g.xoff, 1, h->stream));
    } else {
      HIPCHK(h, (hipError_t)launch_box_move(g, h->stage, es, sub, h->g.px, h->g.pxy, h->g.xoff, 0, h->stream));
      HIPCHK(h, hipMemcpyAsync((char *)host_dst + (size_t)k0 * plane, h->stage, bytes,
                               hipMemcpyDeviceToHost, h->stream));
    }
    // the staging buffer is reused by the next run
    HIPCHK(h, hipStreamSynchronize(h->stream));
  }
  return FDTD_OK;
}

// a field component to / from the host as fp32, whatever the storage
int field_io(fdtd_handle *h, int comp, const float *src, float *dst) {
  if (h->storage == FDTD_STORE_FP32)
    return copy_box(h, comp, h->fld[comp], src, dst, 4);
  long long ext[3];
  alloc_extent(h, comp, ext);
  const size_t n = (size_t)(ext[0] * ext[1] * ext[2]);
  std::vector<__half> tmp(n);
  const float s = h->g.fs[comp], is = h->g.fi[comp];
  if (src) {
    for (size_t t = 0; t < n; ++t) tmp[t] = __float2half_rn(src[t] * is);
    return copy_box(h, comp, h->fld[comp], tmp.data(), nullptr, 2);
  }
  int rc = copy_box(h, comp, h->fld[comp], nullptr, tmp.data(), 2);
  if (rc) return rc;
  for (size_t t = 0; t < n; ++t) dst[t] = __half2float(tmp[t]) * s;
  return FDTD_OK;
}

// stepping calls on a grid that belongs to a batch are refused: the batch's
// captured step replays over the grid's buffers and reads its step counter
int refuse_batched(fdtd_handle *h) {
  return h->batch ? fail(h, FDTD_ERR_STATE, "grid is admitted to a batch: retire it first")
                  : FDTD_OK;
}

// room in the probe buffer for `more` samples beyond those buffered since the
// last read; the probe kernels stop writing at the buffer's end, so a call
// that would overflow it is refused instead of dropping samples
int probe_room(fdtd_handle *h, long long more) {
  if (!h->nprobe || h->pbuffered + more <= h->pcap) return FDTD_OK;
  return fail(h, FDTD_ERR_ARG,
              "probe buffer holds %d steps, %lld buffered since the last read, %lld more asked",
              h->pcap, h->pbuffered, more);
}

int check_ready(fdtd_handle *h) {
  if (h->nmedia < 0) return fail(h, FDTD_ERR_STATE, "fdtd_set_media not called");
  if (h->media_max > h->nmedia)
    return fail(h, FDTD_ERR_STATE, "media index %d written but only %d media set", h->media_max, h->nmedia);
  for (int c = 0; c < 6; ++c)
    if (!h->did[c]) return fail(h, FDTD_ERR_STATE, "fdtd_set_spacing not called");
  return FDTD_OK;
}

// blocks of 64 x rows threads march z in chunks of zchunk planes
void set_tiles(fdtd_handle *h, int rows, int zchunk, int xcells) {
  const DevGrid &g = h->g;
  h->tile_zchunk = zchunk;
  for (int v = 0; v < 2; ++v) {
    TileMap &tm = v ? h->tm4 : h->tm;
    const int w = v ? 4 * kWave : kWave;
    tm.bdy = rows;
    tm.tx = (g.nx + 1 + w - 1) / w;
    tm.ty = (g.ny + 1 + rows - 1) / rows;
    tm.kc = std::min(zchunk, g.nz + 1);
    tm.tz = (g.nz + 1 + tm.kc - 1) / tm.kc;
    tm.nblk = tm.tx * tm.ty * tm.tz;
    tm.remap = 1;
    tm.vec = xcells;
  }
}

static void drop_multi_graph(fdtd_handle *h) {
  if (h->mexec) hipGraphExecDestroy(h->mexec);
  if (h->mgraph) hipGraphDestroy(h->mgraph);
  h->mexec = nullptr;
  h->mgraph = nullptr;
}

void invalidate_graph(fdtd_handle *h) {
  ++h->version;
  drop_multi_graph(h);
  if (h->gexec) hipGraphExecDestroy(h->gexec);
  if (h->graph) hipGraphDestroy(h->graph);
  if (h->pexec) hipGraphExecDestroy(h->pexec);
  if (h->pgraph) hipGraphDestroy(h->pgraph);
  h->gexec = nullptr;
  h->graph = nullptr;
  h->pexec = nullptr;
  h->pgraph = nullptr;
  h->graph_ok = false;
  h->step_ready = false;
  h->tmed_ready = false;
}

// the medium of each tile of the one-cell pass map (or -1): the passes then
// skip the media loads of single-medium tiles.  Built outside any capture,
// before the passes it serves are launched or recorded.
int build_tile_media(fdtd_handle *h) {
  if (h->tmed_ready) return FDTD_OK;
  if (h->tm.nblk > h->tmed_cap) {
    if (h->tmed) hipFree(h->tmed);
    h->tmed = nullptr;
    h->tmed_cap = 0;
    HIPCHK(h, hipMalloc(&h->tmed, (size_t)h->tm.nblk * sizeof(int)));
    h->tmed_cap = h->tm.nblk;
  }
  const int e = h->mode == FDTD_MODE_EXACT ? exact::launch_pass_tile_media(h->g, h->tm, h->tmed, h->stream)
                                           : fast::launch_pass_tile_media(h->g, h->tm, h->tmed, h->stream);
  if (e) return fail(h, FDTD_ERR_HIP, "tile media launch failed: %s", hipGetErrorString((hipError_t)e));
  h->tmed_ready = true;
  return FDTD_OK;
}

// ---- compact pole state --------------------------------------------------
// One field's pole state (side 0: electric media on E, 1: magnetic media on
// H): per-medium coefficients, and per component a dense box over its
// dispersive edges.
static int build_ade_side(fdtd_handle *h, int side) {
  sfd_host::AdeIn &in = h->adein[side];
  AdeSet &ad = h->g.ade[side];
  ad = AdeSet{};
  for (int c = 0; c < 3; ++c) { in.nedges[c] = 0; in.jbytes[c] = 0; }
  if (in.ndisp == 0) return FDTD_OK;
  const int nm = h->nmedia;
  std::vector<int32_t> slot(nm + 1, -1);
  for (int d = 0; d < in.ndisp; ++d) {
    if (in.dmed[d] < 1 || in.dmed[d] > nm)
      return fail(h, FDTD_ERR_ARG, "dispersive medium %d out of range", in.dmed[d]);
    slot[in.dmed[d]] = d;
  }
  int32_t *dslot = nullptr;
  int rc = upload(h, &dslot, slot.data(), slot.size(), &h->ade_allocs);
  if (rc) return rc;
  ad.dslot = dslot;
  // coefficients
  std::vector<int32_t> anp(in.ndisp);
  std::vector<float> cf((size_t)in.ndisp * kPMax * 6, 0.0f);
  std::vector<double> cd((size_t)in.ndisp * kPMax * 6, 0.0);
  for (int d = 0; d < in.ndisp; ++d) {
    const int np = in.dnp[d];
    const double *K = &in.dk[(size_t)d * kPMax * 2], *B = &in.db[(size_t)d * kPMax * 2],
                 *G = &in.dg3[(size_t)d * kPMax * 2];
    for (int q = 0; q < np; ++q) {
      double *o = &cd[((size_t)d * kPMax + q) * 6];
      o[0] = K[2 * q]; o[1] = K[2 * q + 1]; o[2] = B[2 * q]; o[3] = B[2 * q + 1];
      o[4] = G[2 * q]; o[5] = G[2 * q + 1];
    }
    if (h->mode == FDTD_MODE_EXACT) { anp[d] = np; continue; }
    // FAST: fold complex-conjugate pairs (their states stay conjugate, so the
    // pair contributes 2 Re(G3 J) of one state); beta is pre-divided by dt.
    std::vector<int> used(np, 0);
    int m = 0;
    for (int q = 0; q < np; ++q) {
      if (used[q]) continue;
      used[q] = 1;
      double w = 1.0;
      for (int r = q + 1; r < np; ++r)
        if (!used[r] && K[2 * r] == K[2 * q] && K[2 * r + 1] == -K[2 * q + 1] &&
            B[2 * r] == B[2 * q] && B[2 * r + 1] == -B[2 * q + 1] &&
            G[2 * r] == G[2 * q] && G[2 * r + 1] == -G[2 * q + 1] && K[2 * q + 1] != 0.0) {
          used[r] = 1;
          w = 2.0;
          break;
        }
      float *o = &cf[((size_t)d * kPMax + m) * 6];
      o[0] = (float)K[2 * q]; o[1] = (float)K[2 * q + 1];
      o[2] = (float)(B[2 * q] / h->dt); o[3] = (float)(B[2 * q + 1] / h->dt);
      o[4] = (float)(w * G[2 * q]); o[5] = (float)(w * G[2 * q + 1]);
      ++m;
    }
    anp[d] = m;
  }
  int32_t *danp = nullptr;
  if ((rc = upload(h, &danp, anp.data(), anp.size(), &h->ade_allocs))) return rc;
  ad.anp = danp;
  if (h->mode == FDTD_MODE_EXACT) {
    double *p = nullptr;
    if ((rc = upload(h, &p, cd.data(), cd.size(), &h->ade_allocs))) return rc;
    ad.aded = p;
  } else {
    float *p = nullptr;
    if ((rc = upload(h, &p, cf.data(), cf.size(), &h->ade_allocs))) return rc;
    ad.adef = p;
  }
  ad.ndisp = in.ndisp;
  int maxp = 0;
  for (int d = 0; d < in.ndisp; ++d) maxp = std::max(maxp, anp[d]);
  ad.maxp = std::max(maxp, 1);
  // dense state box per component over its dispersive edges (the bounding
  // box kernel reads the slot table through the grid descriptor)
  int32_t *bb = nullptr;
  if ((rc = dalloc(h, (void **)&bb, 8 * 4, &h->ade_allocs))) return rc;
  for (int c = 0; c < 3; ++c) {
    const int32_t init[7] = {0x7fffffff, 0x7fffffff, 0x7fffffff, (int32_t)0x80000000,
                             (int32_t)0x80000000, (int32_t)0x80000000, 0};
    HIPCHK(h, hipMemcpy(bb, init, sizeof init, hipMemcpyHostToDevice));
    HIPCHK(h, (hipError_t)launch_disp_bbox(h->g, side, c, bb, h->stream));
    int32_t box[7];
    HIPCHK(h, hipMemcpyAsync(box, bb, sizeof box, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    in.nedges[c] = box[6];
    long long total = 0;
    if (box[0] <= box[3]) {
      for (int a = 0; a < 3; ++a) {
        ad.abox[c][a] = box[a];
        ad.abox[c][3 + a] = box[3 + a] - box[a] + 1;
      }
      total = (long long)ad.abox[c][3] * ad.abox[c][4] * ad.abox[c][5];
    } else {
      for (int a = 0; a < 6; ++a) ad.abox[c][a] = 0;  // empty box: no edge inside
    }
    ad.nstate[c] = total;
    float *ep = nullptr;
    void *js = nullptr;
    const size_t ns = (size_t)std::max(total, 1LL);
    const size_t jbytes = ns * std::max(maxp, 1) * (h->mode == FDTD_MODE_EXACT ? 16 : 8);
    if ((rc = dalloc(h, (void **)&ep, ns * 4, &h->ade_allocs))) return rc;
    if ((rc = dalloc(h, &js, jbytes, &h->ade_allocs))) return rc;
    in.jbytes[c] = jbytes;
    if ((rc = dzero(h, ep, ns * 4))) return rc;
    if ((rc = dzero(h, js, jbytes))) return rc;
    ad.prev[c] = ep;
    ad.jst[c] = js;
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  // no dispersive edge anywhere: the passes run without this pole stage
  ad.has = (in.nedges[0] + in.nedges[1] + in.nedges[2]) > 0 ? 1 : 0;
  return FDTD_OK;
}

int build_ade(fdtd_handle *h) {
  if (h->ade_ready) return FDTD_OK;
  for (void *p : h->ade_allocs) hipFree(p);
  h->ade_allocs.clear();
  for (int side = 0; side < 2; ++side) {
    const int rc = build_ade_side(h, side);
    if (rc) return rc;
  }
  h->ade_ready = true;
  return FDTD_OK;
}

// ---- first-order Mur faces -------------------------------------------------
// Runs in the reference's face order (left, right, down, up, back, front,
// bordersmur.F90:1129-1383), each face's two tangential H components; the
// components' sweeps give the face extents (InitMURBorders, :113-160).
static int build_mur(fdtd_handle *h) {
  if (h->mur_ready) return FDTD_OK;
  for (void *p : h->mur_allocs) hipFree(p);
  h->mur_allocs.clear();
  h->mur_runs.clear();
  h->dmur = nullptr;
  h->mur_total = 0;
  bool any = false;
  for (int f = 0; f < 6; ++f) any = any || h->mur_faces[f];
  if (!any) { h->mur_ready = true; return FDTD_OK; }
  if ((int)h->cab1h.size() != 6 * (h->nmedia + 1))
    return fail(h, FDTD_ERR_STATE, "Mur coefficients do not match the media table");
  float *dcab = nullptr;
  int rc = upload(h, &dcab, h->cab1h.data(), h->cab1h.size(), &h->mur_allocs);
  if (rc) return rc;
  static const int order[6] = {2, 3, 4, 5, 0, 1};
  static const int comps[6][2] = {{5, 4}, {5, 4}, {3, 5}, {3, 5}, {4, 3}, {4, 3}};
  const int n[3] = {h->g.nx, h->g.ny, h->g.nz};
  for (int o = 0; o < 6; ++o) {
    const int f = order[o];
    if (!h->mur_faces[f]) continue;
    const int a = f / 2, hi = f & 1;
    for (int t = 0; t < 2; ++t) {
      MurRun r{};
      r.comp = comps[f][t];
      const int c = r.comp - 3;
      // H component c sweeps [0, n] along its own axis, [0, n-1] across
      auto last = [&](int ax) { return ax == c ? n[ax] : n[ax] - 1; };
      r.axis = a;
      r.ghost = hi ? last(a) + 1 : -1;
      r.inner = hi ? last(a) : 0;
      r.ua = a == 0 ? 1 : 0;
      r.va = a == 2 ? 1 : 2;
      r.nu = last(r.ua) + 1;
      r.nv = last(r.va) + 1;
      r.start = h->mur_total;
      r.cab1 = dcab + (size_t)f * (h->nmedia + 1);
      const size_t bytes = (size_t)2 * r.nu * r.nv * 4;
      if ((rc = dalloc(h, (void **)&r.past, bytes, &h->mur_allocs))) return rc;
      if ((rc = dzero(h, r.past, bytes))) return rc;
      h->mur_total += (long long)r.nu * r.nv;
      h->mur_runs.push_back(r);
    }
  }
  if ((rc = upload(h, &h->dmur, h->mur_runs.data(), h->mur_runs.size(), &h->mur_allocs)))
    return rc;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->mur_ready = true;
  return FDTD_OK;
}

int run_mur(fdtd_handle *h) {
  if (h->mur_runs.empty()) return FDTD_OK;
  if (launch_mur(h->g, h->dmur, (int)h->mur_runs.size(), h->mur_total,
                 h->mode == FDTD_MODE_EXACT, h->stream))
    return fail(h, FDTD_ERR_HIP, "Mur launch failed");
  return FDTD_OK;
}

// ---- one launch per step --------------------------------------------------
// k_step reads the state of step n from one copy and writes step n+1 to the
// other; fdtd_run advances pairs of steps A -> B -> A, so between calls the
// state is always in A and every other entry point sees it there.
bool step_supported(const fdtd_handle *h, int rows) {
  // lanes address a plane by a 32-bit byte offset; a tile's end plane fits
  // the low bits of its k1 word (a single-medium tile's medium sits above)
  if (h->g.fp16 || h->nmedia + 1 > 4096 || h->g.pxy * 4 >= (1LL << 31) ||
      h->g.nz + 1 > (int)kTileK1Mask)
    return false;
  if ((int)h->pw_faces[0].size() > kStepFaces || (int)h->pw_faces[1].size() > kStepFaces)
    return false;
  DevGrid g = h->g;
  const int npw = (int)(h->pw_faces[0].size() + h->pw_faces[1].size());
  const int kc = std::min(h->step_kc, h->g.nz + 1);
  const size_t lds = h->mode == FDTD_MODE_EXACT ? exact::step_lds_bytes(g, rows, npw, kc)
                                                : fast::step_lds_bytes(g, rows, npw, kc);
  return lds <= 48 * 1024;
}

// rows per block of tile class c: the rows asked for, if they cut the step's
// tiles into whole bands, else the step's.  By default the class with CPML,
// poles and sources runs 4-row blocks under 16-row tiles: its 16-row kernel
// needs more registers than 1024-thread blocks leave a lane and spills.
int class_rows(const fdtd_handle *h, int c) {
  const int R = h->step_rows;
  if (R <= 0) return 0;
  if (c == kPlanVec) {
    // the four-cell kernel: 4, 8 or 16 rows (the step's by default)
    const int q = h->step_creq[c] ? h->step_creq[c] : R;
    return (q == 4 || q == 8 || q == 16) && q <= R ? q : R;
  }
  if (c >= kPlanHalf) {
    // half-width tiles are banded on their own (2E - 1 rows for E waves), so
    // any supported block fits.  Defaults measured on the 1024^3 grid under
    // 16-row steps: CPML + poles + sources (more than the 128 registers a
    // lane has in 16-wave blocks) on 4 waves, CPML + sources on 8
    const int b = c - kPlanHalf;
    const int q = h->step_creq[c] ? h->step_creq[c]
                  : b == 7          ? std::min(R, 4)
                  : b == 5 && R > 8 ? 8
                                    : R;
    return q <= R ? q : R;
  }
  // CPML-only and CPML + poles + sources tiles in 4-row blocks under 16-row
  // steps (measured the faster on the 1024^3 grid)
  const int q = h->step_creq[c] ? h->step_creq[c] : ((c == 7 || c == 1) && R == 16 ? 4 : R);
  return q > 1 && q <= R && (R - 1) % (q - 1) == 0 ? q : R;
}

int build_step(fdtd_handle *h) {
  if (h->step_ready) return FDTD_OK;
  const int R = h->step_rows;
  if (!step_supported(h, R))
    return fail(h, FDTD_ERR_STATE, "one-launch steps need fp32 storage, small media tables and at most 64 TF/SF faces per half");
  // layout of the state: fields, psi arrays, pole state of both fields
  const size_t ne = field_elems(h);
  std::vector<size_t> sizes(6, ne * 4);
  for (int s = 0; s < 12; ++s) sizes.push_back(h->g.psi[s] ? (size_t)h->psilen[s] * 4 : 0);
  for (int side = 0; side < 2; ++side)
    for (int c = 0; c < 3; ++c) {
      const AdeSet &a = h->g.ade[side];
      sizes.push_back(a.prev[c] ? (size_t)std::max(a.nstate[c], 1LL) * 4 : 0);
      sizes.push_back(a.jst[c] ? h->adein[side].jbytes[c] : 0);
    }
  // 8-bit copies of the media indices: a sixth less traffic per cell
  const bool narrow = h->nmedia + 1 <= 256;
  for (int c = 0; c < 6; ++c) sizes.push_back(narrow ? ne : 0);
  int rc;
  std::vector<void *> ptr(sizes.size(), nullptr);
  if (sizes != h->step_sizes) {
    for (void *p : h->step_allocs) { hipFree(p); }
    h->step_allocs.clear();
    h->step_sizes.clear();
    for (size_t q = 0; q < sizes.size(); ++q) {
      if (!sizes[q]) continue;
      if ((rc = dalloc(h, &ptr[q], sizes[q], &h->step_allocs)) || (rc = dzero(h, ptr[q], sizes[q])))
        return rc;
    }
    h->step_sizes = sizes;
    h->step_stale = true;
    for (int c = 0; c < 6; ++c) h->mi8[c] = (uint8_t *)ptr[30 + c];
    h->gB = h->g;
    for (int c = 0; c < 6; ++c) h->gB.f[c] = (float *)ptr[c];
    for (int s = 0; s < 12; ++s) h->gB.psi[s] = (float *)ptr[6 + s];
    for (int side = 0; side < 2; ++side)
      for (int c = 0; c < 3; ++c) {
        h->gB.ade[side].prev[c] = (float *)ptr[18 + side * 6 + c * 2];
        h->gB.ade[side].jst[c] = ptr[18 + side * 6 + c * 2 + 1];
      }
  }
  // the 8-bit media copies are refreshed from the int16 arrays at every
  // set-up (the media may have changed since)
  if (narrow)
    for (int c = 0; c < 6; ++c) {
      if (!h->mi8[c]) return fail(h, FDTD_ERR_STATE, "8-bit media copy missing");
      if (launch_media_narrow(h->g.mi[c], h->mi8[c], (long long)ne, h->stream))
        return fail(h, FDTD_ERR_HIP, "media narrowing failed");
    }
  // everything but the state pointers follows A
  DevGrid gb = h->g;
  for (int c = 0; c < 6; ++c) gb.f[c] = h->gB.f[c];
  for (int s = 0; s < 12; ++s) gb.psi[s] = h->gB.psi[s];
  for (int side = 0; side < 2; ++side)
    for (int c = 0; c < 3; ++c) {
      gb.ade[side].prev[c] = h->gB.ade[side].prev[c];
      gb.ade[side].jst[c] = h->gB.ade[side].jst[c];
    }
  h->gB = gb;
  for (void *p : h->step_small) hipFree(p);
  h->step_small.clear();
  StepDev sd{};
  TileMap &tm = sd.tm;
  tm.bdy = R;
  tm.vec = 1;
  tm.kc = std::min(h->step_kc, h->g.nz + 1);
  tm.remap = 1;
  StepSrc src{};
  // the faces of both halves in one list, each with its slice of the
  // incident-field buffer
  {
    std::vector<PwFace> all;
    long long cells = 0, maxc = 0;
    for (int half = 0; half < 2; ++half)
      for (PwFace f : h->pw_faces[half]) {
        const int ua = f.axis == 0 ? 1 : 0, va = f.axis == 2 ? 1 : 2;
        const long long nc = (long long)(f.hi[ua] - f.lo[ua] + 1) * (f.hi[va] - f.lo[va] + 1);
        f.inc = cells;
        cells += std::max(nc, 0LL);
        maxc = std::max(maxc, nc);
        all.push_back(f);
      }
    h->step_nf0 = (int)h->pw_faces[0].size();
    h->step_nf = (int)all.size();
    h->step_maxcells = maxc;
    h->dsfaces = nullptr;
    h->dinc = nullptr;
    if (!all.empty()) {
      if ((rc = upload(h, &h->dsfaces, all.data(), all.size(), &h->step_small))) return rc;
      if ((rc = dalloc(h, (void **)&h->dinc, (size_t)std::max(cells, 1LL) * 4, &h->step_small)))
        return rc;
    }
    sd.inc = h->dinc;
  }
  for (int half = 0; half < 2; ++half) {
    const std::vector<PwFace> &F = h->pw_faces[half];
    sd.npw[half] = src.npw[half] = (int)F.size();
    src.pwf[half] = F.empty() ? nullptr : h->dsfaces + (half ? h->step_nf0 : 0);
    const std::vector<NodalCell> &C = h->hcells[half];
    sd.ncell[half] = src.ncell[half] = (int)C.size();
    src.cells[half] = h->dcells[half];
    int *b = sd.cellbox[half];
    b[0] = b[1] = b[2] = 0x7fffffff;
    b[3] = b[4] = b[5] = -0x7fffffff;
    for (const NodalCell &c : C) {
      const long long k = c.p / h->g.pxy - 1, r = c.p % h->g.pxy;
      const long long j = r / h->g.px - 1, i = r % h->g.px - h->g.xoff;
      const long long q[3] = {i, j, k};
      for (int a = 0; a < 3; ++a) {
        b[a] = (int)std::min<long long>(b[a], q[a]);
        b[3 + a] = (int)std::max<long long>(b[3 + a], q[a]);
      }
    }
  }
  // the tile plan (fdtd_plan.cpp): tiles cut at the CPML slabs, the TF/SF
  // faces, the nodal cells and the pole-state boxes, one launch per class
  {
    PlanGeom pg;
    pg.n[0] = h->g.nx; pg.n[1] = h->g.ny; pg.n[2] = h->g.nz;
    for (int a = 0; a < 3; ++a)
      for (int s2 = 0; s2 < 2; ++s2) pg.L[a][s2] = h->g.has_cpml ? h->g.L[a][s2] : 0;
    pg.R = R;
    pg.kc = tm.kc;
    for (int c = 0; c < 8; ++c) {
      pg.rows[c] = class_rows(h, c);
      pg.hrows[c] = class_rows(h, kPlanHalf + c);
    }
    pg.vrows = class_rows(h, kPlanVec);
    // the four-cell kernel's LDS (tables and its row exchange) must fit
    {
      DevGrid g = h->g;
      const size_t l4 = h->mode == FDTD_MODE_EXACT ? exact::step4_lds_bytes(g, class_rows(h, kPlanVec))
                                                   : fast::step4_lds_bytes(g, class_rows(h, kPlanVec));
      pg.vec = h->step_vec && l4 <= (size_t)h->lds_max;
    }
    for (int side = 0; side < 2; ++side) {
      const AdeSet &A = h->g.ade[side];
      if (!A.has) continue;
      for (int c = 0; c < 3; ++c) {
        const int *b = A.abox[c];
        if (b[3] <= 0 || b[4] <= 0 || b[5] <= 0) continue;
        pg.boxes.push_back({{b[0], b[1], b[2]}, {b[0] + b[3] - 1, b[1] + b[4] - 1, b[2] + b[5] - 1}, 2,
                            3 * side + c});
      }
    }
    for (int half = 0; half < 2; ++half) {
      for (const PwFace &f : h->pw_faces[half]) {
        PlanBox b;
        for (int a = 0; a < 3; ++a) { b.lo[a] = f.lo[a]; b.hi[a] = f.hi[a]; }
        b.lo[f.axis] = b.hi[f.axis] = f.fixed;
        b.kind = 4;
        pg.boxes.push_back(b);
      }
      if (sd.ncell[half] > 0) {
        const int *c = sd.cellbox[half];
        pg.boxes.push_back({{c[0], c[1], c[2]}, {c[3], c[4], c[5]}, 4});
      }
    }
    StepPlan plan;
    std::string perr;
    if (plan_step(pg, &plan, &perr)) return fail(h, FDTD_ERR_ARG, "%s", perr.c_str());
    std::vector<StepTile> all;
    h->step_cls.clear(); h->step_cnt.clear(); h->step_off.clear(); h->step_cells.clear();
    h->step_bytes.clear();
    // compulsory bytes of a class launch: each field read and written once
    // and each media index read once per owned cell, 8 B per psi value (read
    // and write), per pole-state cell the stored poles (complex64 folded
    // pairs in FAST, complex128 in EXACT; read and write) and, in tiles with
    // sources, the previous field (8 B; the other tiles take it from the
    // field and leave its store to the settling pass)
    const double cb = 48.0 + 6.0 * (narrow ? 1.0 : 2.0);
    const double per = h->mode == FDTD_MODE_EXACT ? 32.0 : 16.0;
    h->step_crows.clear();
    for (int c = 0; c < kPlanClasses; ++c) {
      if (plan.tiles[c].empty()) continue;
      if (plan.tiles[c].size() > 0x7fffffffULL)
        return fail(h, FDTD_ERR_ARG, "step tiling gives too many blocks");
      h->step_cls.push_back(c);
      h->step_cells.push_back(plan.owned[c]);
      double by = cb * plan.owned[c] + 8.0 * plan.psi[c];
      const int bits = c >= kPlanHalf ? c - kPlanHalf : c;
      const double prevb = c != kPlanVec && (bits & 4) ? 8.0 : 0.0;
      for (int side = 0; side < 2; ++side)
        if (h->g.ade[side].has)
          by += plan.disp[c][side] * (prevb + per * std::max(h->g.ade[side].maxp, 1));
      h->step_bytes.push_back(by);
      h->step_cnt.push_back((int)plan.tiles[c].size());
      h->step_off.push_back((int)all.size());
      h->step_crows.push_back(class_rows(h, c));
      all.insert(all.end(), plan.tiles[c].begin(), plan.tiles[c].end());
    }
    h->dtiles = nullptr;
    if (!all.empty() && (rc = upload(h, &h->dtiles, all.data(), all.size(), &h->step_small)))
      return rc;
    // four-cell tiles whose cells all hold one medium (every component, the
    // halo included) skip the media loads: they lead their class's tiles,
    // the medium packed above k1's plane bits, and their compulsory bytes
    // lose the media indices
    h->step_uni.assign(h->step_cls.size(), 0);
    bool packed = false;
    for (size_t q = 0; q < h->step_cls.size(); ++q) {
      if (h->step_cls[q] != kPlanVec || h->nmedia + 1 > (1 << (32 - kTileMediumShift)))
        continue;
      const int n = h->step_cnt[q], off = h->step_off[q];
      int *dm = nullptr;
      if ((rc = dalloc(h, (void **)&dm, (size_t)n * 4, &h->step_small))) return rc;
      if (launch_tile_media(h->g, h->dtiles + off, n, dm, h->stream))
        return fail(h, FDTD_ERR_HIP, "tile media launch failed");
      std::vector<int> m((size_t)n);
      HIPCHK(h, hipMemcpyAsync(m.data(), dm, (size_t)n * 4, hipMemcpyDeviceToHost, h->stream));
      HIPCHK(h, hipStreamSynchronize(h->stream));
      std::vector<StepTile> uni, mixed;
      double uowned = 0.0;
      for (int t = 0; t < n; ++t) {
        StepTile s = all[(size_t)off + (size_t)t];
        if (m[(size_t)t] < 0) { mixed.push_back(s); continue; }
        const int x0 = s.x & 0xfffff, wx = s.x >> 20, y0 = s.y & 0xfffff, wy = s.y >> 20;
        uowned += (double)std::max(0, std::min(x0 + wx, h->g.nx) - x0) *
                  std::max(0, std::min(y0 + wy, h->g.ny) - y0) *
                  std::max(0, std::min(s.k1, h->g.nz) - s.k0);
        s.k1 = (int)((unsigned)s.k1 | ((unsigned)m[(size_t)t] << kTileMediumShift));
        uni.push_back(s);
      }
      std::copy(uni.begin(), uni.end(), all.begin() + off);
      std::copy(mixed.begin(), mixed.end(), all.begin() + off + (long)uni.size());
      h->step_uni[q] = (int)uni.size();
      h->step_bytes[q] -= 6.0 * (narrow ? 1.0 : 2.0) * uowned;
      packed = packed || !uni.empty();
    }
    if (packed) {
      HIPCHK(h, hipMemcpyAsync(h->dtiles, all.data(), all.size() * sizeof(StepTile),
                               hipMemcpyHostToDevice, h->stream));
      HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    h->step_tiles = all;
  }
  src.waves = h->dwaves;
  src.phys = h->dphys;
  src.pstride = h->pstride;
  src.tiempo = h->dtiempo;
  src.dt = h->dt;
  src.pwG[0] = h->nmedia >= 1 ? h->g2h[1] : 0.0f;
  src.pwG[1] = h->nmedia >= 1 ? h->gm2h[1] : 0.0f;
  src.values = h->dvalues;
  src.nsrc = (int)h->srcs.size();
  StepSrc *dsrc = nullptr;
  if ((rc = upload(h, &dsrc, &src, 1, &h->step_small))) return rc;
  sd.src = dsrc;
  h->dstepsrc = dsrc;
  sd.narrow = narrow ? 1 : 0;
  for (int c = 0; c < 6; ++c) sd.mi8[c] = narrow ? h->mi8[c] : nullptr;
  sd.pwG[0] = src.pwG[0];
  sd.pwG[1] = src.pwG[1];
  sd.dn = h->dstep;
  // CPML and pole state of each direction, read by the slab and dispersive
  // cells
  for (int dir = 0; dir < 2; ++dir) {
    const DevGrid &gi = dir ? h->gB : h->g;
    const DevGrid &go = dir ? h->g : h->gB;
    StepAux a{};
    for (int c = 0; c < 6; ++c) {
      a.fi[c] = gi.f[c]; a.fo[c] = go.f[c]; a.mi[c] = gi.mi[c];
      a.mi8[c] = narrow ? h->mi8[c] : nullptr;
    }
    for (int s = 0; s < 12; ++s) { a.psi_i[s] = gi.psi[s]; a.psi_o[s] = go.psi[s]; }
    for (int ax = 0; ax < 3; ++ax) {
      a.pbe[ax] = gi.pbe[ax]; a.pce[ax] = gi.pce[ax];
      a.pbm[ax] = gi.pbm[ax]; a.pcm[ax] = gi.pcm[ax];
      a.L[ax][0] = gi.L[ax][0]; a.L[ax][1] = gi.L[ax][1];
    }
    for (int side = 0; side < 2; ++side) { a.ai[side] = gi.ade[side]; a.ao[side] = go.ade[side]; }
    a.nx = gi.nx; a.ny = gi.ny; a.nz = gi.nz;
    a.dt = gi.dt;
    StepAux *da = nullptr;
    if ((rc = upload(h, &da, &a, 1, &h->step_small))) return rc;
    h->sd[dir] = sd;
    h->sd[dir].gi = gi;
    h->sd[dir].go = go;
    h->sd[dir].aux = da;
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->step_ready = true;
  return FDTD_OK;
}

// cells outside the sweeps are never written by a step, so B's must equal
// A's: refreshed after the fields were set from outside
int sync_b(fdtd_handle *h) {
  if (!h->step_stale) return FDTD_OK;
  const size_t ne = field_elems(h);
  for (int c = 0; c < 6; ++c)
    HIPCHK(h, hipMemcpyAsync(h->gB.f[c], h->fld[c], ne * 4, hipMemcpyDeviceToDevice, h->stream));
  h->step_stale = false;
  return FDTD_OK;
}

// one step of the pair, from sd[dir].gi to sd[dir].go, then the H-side
// boundary, observation and step counter on the written copy
// the step kernels of one direction: one launch per tile class
// The class kernels of one step direction, one after another on the
// handle's stream.  (They read the same state and write disjoint cells, but
// running the small classes on side streams beside the interior one was
// measured slower: they compete for the same bandwidth and caches.)
// the descriptor of class launch q in direction dir
static StepDev class_sd(const fdtd_handle *h, int dir, size_t q) {
  StepDev sd = h->sd[dir];
  sd.tiles = h->dtiles + h->step_off[q];
  sd.tm.nblk = h->step_cnt[q];
  return sd;
}

static int launch_class(const fdtd_handle *h, const StepDev &sd, size_t q) {
  const int nuni = h->step_uni[q];
  return h->mode == FDTD_MODE_EXACT
             ? exact::launch_step(sd, h->step_crows[q], h->step_cls[q], nuni, h->stream)
             : fast::launch_step(sd, h->step_crows[q], h->step_cls[q], nuni, h->stream);
}

static int launch_step_kernels(fdtd_handle *h, int dir, bool exact_prev = false) {
  if (h->step_nf > 0 &&
      launch_pw_incident(h->dsfaces, h->step_nf0, h->step_nf, h->step_maxcells, h->dstepsrc,
                         h->dstep, h->dinc, h->stream))
    return fail(h, FDTD_ERR_HIP, "incident-field launch failed");
  for (size_t q = 0; q < h->step_cls.size(); ++q) {
    StepDev sd = class_sd(h, dir, q);
    sd.prev_from_field = exact_prev ? 0 : 1;
    sd.lazy_prev = 1;
    const int e = launch_class(h, sd, q);
    if (e) return fail(h, FDTD_ERR_HIP, "step launch failed: %s", hipGetErrorString((hipError_t)e));
  }
  return FDTD_OK;
}

// After steps that left the stored previous field of the pole classes
// without sources to this pass: set it from the fields of the state (A), so
// that every other reader (the two passes, checkpoints, the next plan, a
// batch) finds it exact.  Called before any entry point returns.
static int settle_prev(fdtd_handle *h) {
  if (!h->prev_lazy) return FDTD_OK;
  for (size_t q = 0; q < h->step_cls.size(); ++q) {
    const int c = h->step_cls[q] >= kPlanHalf ? h->step_cls[q] - kPlanHalf : h->step_cls[q];
    if (h->step_cls[q] == kPlanVec || !(c & 2) || (c & 4)) continue;
    if (launch_prev_settle(h->g, h->dtiles + h->step_off[q], h->step_cnt[q], h->stream))
      return fail(h, FDTD_ERR_HIP, "previous-field settling launch failed");
  }
  h->prev_lazy = false;
  return FDTD_OK;
}

static int run_step(fdtd_handle *h, int dir, bool exact_prev = false) {
  const StepDev &sd = h->sd[dir];
  int e = launch_step_kernels(h, dir, exact_prev);
  if (e) return e;
  if (!h->mur_runs.empty() &&
      launch_mur(sd.go, h->dmur, (int)h->mur_runs.size(), h->mur_total,
                 h->mode == FDTD_MODE_EXACT, h->stream))
    return fail(h, FDTD_ERR_HIP, "Mur launch failed");
  if (h->nprobe &&
      launch_probes(sd.go, h->dprobes, h->nprobe, h->dpbuf, h->dpcount, h->pcap, h->stream))
    return fail(h, FDTD_ERR_HIP, "probe launch failed");
  for (const DftDev &d : h->dfts)
    if (launch_dft(sd.go, d, 0, h->dstep, h->stream))
      return fail(h, FDTD_ERR_HIP, "DFT launch failed");
  if (launch_tick(h->dstep, h->stream)) return fail(h, FDTD_ERR_HIP, "tick failed");
  return FDTD_OK;
}

// ---- nodal sources --------------------------------------------------------
int build_nodal(fdtd_handle *h) {
  if (h->nodal_ready) return FDTD_OK;
  if (h->srcs.empty()) { h->nodal_ready = true; return FDTD_OK; }
  if (h->tiempo.empty()) return fail(h, FDTD_ERR_STATE, "fdtd_set_time not called");
  std::vector<NodalCell> cells[2];
  for (int half = 0; half < 2; ++half) {
    for (int c = 3 * half; c < 3 * half + 3; ++c)
      for (int hard = 1; hard >= 0; --hard)
        for (size_t s = 0; s < h->srcs.size(); ++s) {
          const Source &S = h->srcs[s];
          if (S.comp != c || S.hard != hard) continue;
          for (int k = S.box[4]; k <= S.box[5]; ++k)
            for (int j = S.box[2]; j <= S.box[3]; ++j)
              for (int i = S.box[0]; i <= S.box[1]; ++i) {
                NodalCell nc{};
                nc.p = (long long)(k + 1) * h->g.pxy + (long long)(j + 1) * h->g.px + h->g.xoff + i;
                nc.comp = c;
                nc.hard = hard;
                nc.src = (int)s;
                nc.amp = S.amp;
                // media read by the reference: own matrix, except hard Hy/Hz
                // which read sggMiHx (nodalsources.F90:641,715)
                const int mc = (c >= 4 && hard) ? 3 : c;
                int16_t m = 0;
                HIPCHK(h, hipMemcpy(&m, h->mi[mc] + nc.p, 2, hipMemcpyDeviceToHost));
                const int skipbit = c < 3 ? FDTD_KIND_PEC : FDTD_KIND_PMC;
                nc.skip = (h->kindh[m] & skipbit) ? 1 : 0;
                if (!hard) {
                  const float *Ix = c < 3 ? h->idh_[3].data() : h->idh_[0].data();
                  const float *Iy = c < 3 ? h->idh_[4].data() : h->idh_[1].data();
                  const float *Iz = c < 3 ? h->idh_[5].data() : h->idh_[2].data();
                  const float gg = c < 3 ? h->g2h[m] : h->gm2h[m];
                  float ia, ib;
                  switch (c % 3) {
                    case 0: ia = Iy[j + 1]; ib = Iz[k + 1]; break;
                    case 1: ia = Ix[i + 1]; ib = Iz[k + 1]; break;
                    default: ia = Iy[j + 1]; ib = Ix[i + 1]; break;
                  }
                  nc.coef = gg * ia * ib;
                }
                cells[half].push_back(nc);
              }
        }
  }
  // excitation values per step (real(4) time stamps: E tiempo(n), H
  // tiempo(n)+0.5*dt; nodalsources.F90:426,618)
  const int nsteps = (int)h->tiempo.size() - 1;
  const int ns = (int)h->srcs.size();
  std::vector<float> vals((size_t)(nsteps + 1) * ns);
  for (int n = 0; n <= nsteps; ++n)
    for (int s = 0; s < ns; ++s) {
      const Source &S = h->srcs[s];
      const float t = S.comp < 3 ? (float)h->tiempo[n]
                                 : (float)(h->tiempo[n] + (double)0.5f * h->dt);
      vals[(size_t)n * ns + s] = fdtd_evolucion(t, S.deltaevol, S.numus, S.evol.data());
    }
  int rc;
  for (void *p : h->nodal_allocs) hipFree(p);
  h->nodal_allocs.clear();
  h->dvalues = nullptr;
  h->dcells[0] = h->dcells[1] = nullptr;
  if ((rc = dalloc(h, (void **)&h->dvalues, vals.size() * 4, &h->nodal_allocs))) return rc;
  HIPCHK(h, hipMemcpy(h->dvalues, vals.data(), vals.size() * 4, hipMemcpyHostToDevice));
  h->nvalues_steps = nsteps + 1;
  for (int half = 0; half < 2; ++half) {
    h->ncells[half] = (int)cells[half].size();
    // the reference applies the cells in order; run them serially when two
    // entries touch the same field value
    std::vector<std::pair<long long, int>> key;
    for (auto &c : cells[half]) key.push_back({c.p, c.comp});
    std::sort(key.begin(), key.end());
    h->serial[half] = std::adjacent_find(key.begin(), key.end()) != key.end();
    h->dcells[half] = nullptr;
    h->hcells[half] = cells[half];
    if (!cells[half].empty()) {
      if ((rc = dalloc(h, (void **)&h->dcells[half], cells[half].size() * sizeof(NodalCell),
                       &h->nodal_allocs))) return rc;
      HIPCHK(h, hipMemcpy(h->dcells[half], cells[half].data(),
                          cells[half].size() * sizeof(NodalCell), hipMemcpyHostToDevice));
    }
  }
  h->nodal_ready = true;
  return FDTD_OK;
}

// ---- TF/SF plane waves ------------------------------------------------------
int sweep_last(const fdtd_handle *h, int c, int a) {
  const bool node = c < 3 ? (c != a) : (c - 3 == a);
  return node ? axis_n(h, a) : axis_n(h, a) - 1;
}
// SINPMLSweep of component c along axis a: the sweep without the CPML slabs
int sin_lo(const fdtd_handle *h, int, int a) { return h->g.L[a][0]; }
int sin_hi(const fdtd_handle *h, int c, int a) { return sweep_last(h, c, a) - h->g.L[a][1]; }

// InitPlaneWave (planewaves.F90:72-766) and calc_planewaveconstants
// (:1494-1509) on the host, in real(4) with the reference's expression order;
// the faces of AdvancePlaneWaveE/H (:878-1475) become one PwFace each, in the
// order the reference applies them.
int build_planewaves(fdtd_handle *h) {
  if (h->pw_ready) return FDTD_OK;
  for (void *p : h->pw_allocs) hipFree(p);
  h->pw_allocs.clear();
  h->pw_faces[0].clear();
  h->pw_faces[1].clear();
  h->dwaves = nullptr; h->dphys = nullptr; h->dtiempo = nullptr;
  if (h->pwin.empty()) { h->pw_ready = true; return FDTD_OK; }
  if (h->tiempo.empty()) return fail(h, FDTD_ERR_STATE, "fdtd_set_time not called");
  if (h->line[0].empty()) return fail(h, FDTD_ERR_STATE, "fdtd_set_lines not called");
  if (h->nmedia < 1) return fail(h, FDTD_ERR_STATE, "plane waves need the background medium 1");
  const int n3[3] = {h->g.nx, h->g.ny, h->g.nz};
  const int S = std::max(n3[0], std::max(n3[1], n3[2])) + 3;
  // physical coordinates (:94-164): node axes on the line, cell axes midway
  std::vector<float> phys((size_t)18 * S, 0.0f);
  for (int c = 0; c < 6; ++c)
    for (int a = 0; a < 3; ++a) {
      const bool node = c < 3 ? (c != a) : (c - 3 == a);
      const int top = sweep_last(h, c, a) + (node ? 1 : 0);
      float *v = &phys[(size_t)(c * 3 + a) * S];
      const std::vector<float> &L = h->line[a];
      for (int i = -1; i <= top; ++i)
        v[i + 1] = node ? L[i + 1] : (L[i + 1] + L[i + 2]) * 0.5f;
    }
  std::vector<PwWave> waves(h->pwin.size());
  std::vector<std::vector<float>> evols(h->pwin.size());
  for (size_t q = 0; q < h->pwin.size(); ++q) {
    const PlaneWaveIn &P = h->pwin[q];
    const int *E = P.esq;
    PwWave &w = waves[q];
    std::memset(&w, 0, sizeof w);
    w.nmodes = P.nmodes;
    w.numus = (int)P.evol.size() - 1;
    w.deltaevol = P.deltaevol;
    w.cluz = 1.0f / std::sqrt(P.eps0 * P.mu0);
    const float zvac = std::sqrt(P.mu0 / P.eps0);
    for (int m = 0; m < P.nmodes; ++m) {
      const float px = P.p[m][0], py = P.p[m][1], pz = P.p[m][2];
      const float modulus = std::sqrt(px * px + py * py + std::pow(pz, 2.0f));
      float *pn = w.pn[m];
      pn[0] = px / modulus; pn[1] = py / modulus; pn[2] = pz / modulus;
      w.fpw[0][m] = P.e[m][0]; w.fpw[1][m] = P.e[m][1]; w.fpw[2][m] = P.e[m][2];
      w.fpw[3][m] = (pn[1] * P.e[m][2] - pn[2] * P.e[m][1]) / zvac;
      w.fpw[4][m] = (pn[2] * P.e[m][0] - pn[0] * P.e[m][2]) / zvac;
      w.fpw[5][m] = (pn[0] * P.e[m][1] - pn[1] * P.e[m][0]) / zvac;
      // null-phase corner (:372-407) and box diagonal (:408-411)
      const int hc[3] = {3, 4, 5};
      float c0[3], sq[3];
      for (int a = 0; a < 3; ++a) {
        const int clo = std::max(E[2 * a] - 1, sin_lo(h, hc[a], a));
        const int chi = std::min(E[2 * a + 1] + 1, sin_hi(h, hc[a], a));
        const std::vector<float> &L = h->line[a];
        c0[a] = pn[a] >= 0.0f ? L[clo + 1] : L[chi + 1];
        sq[a] = std::pow(L[clo + 1] - L[chi + 1], 2.0f);
      }
      const float diag = std::sqrt(sq[0] + sq[1] + sq[2]);
      w.dist0[m] = (c0[0] * pn[0] + c0[1] * pn[1] + c0[2] * pn[2]) - P.incert[m] * diag;
    }
    evols[q] = P.evol;
    // illuminated faces (:203-214)
    const bool il[6] = {E[0] >= sin_lo(h, 3, 0) && E[0] <= sin_hi(h, 3, 0),
                        E[1] <= sin_hi(h, 3, 0) && E[1] >= sin_lo(h, 3, 0),
                        E[2] >= sin_lo(h, 4, 1) && E[2] <= sin_hi(h, 4, 1),
                        E[3] <= sin_hi(h, 4, 1) && E[3] >= sin_lo(h, 4, 1),
                        E[4] >= sin_lo(h, 5, 2) && E[4] <= sin_hi(h, 5, 2),
                        E[5] <= sin_hi(h, 5, 2) && E[5] >= sin_lo(h, 5, 2)};
    // free-axis ranges of each E component (:216-295): along its own axis the
    // range stops one short of the box end; H faces copy the paired E face
    int rlo[6][3], rhi[6][3], flo[6][3], fhi[6][3];
    for (int c = 0; c < 3; ++c)
      for (int a = 0; a < 3; ++a) {
        rlo[c][a] = std::max(sin_lo(h, c, a), E[2 * a]);
        rhi[c][a] = std::min(sin_hi(h, c, a), E[2 * a + 1] - (a == c ? 1 : 0));
        flo[c][a] = std::max(sin_lo(h, c, a), E[2 * a]);
        fhi[c][a] = std::min(sin_hi(h, c, a), E[2 * a + 1]);
      }
    // H component -> (paired E component) per face axis: x faces Hy<-Ez,
    // Hz<-Ey; y faces Hz<-Ex, Hx<-Ez; z faces Hx<-Ey, Hy<-Ex
    auto face = [&](int half, int comp, int axis, bool high, int nf, int off, float sign) {
      int ec = comp;
      if (comp >= 3) {
        static const int pair[3][6] = {{0, 0, 0, 0, 2, 1}, {0, 0, 0, 2, 0, 0}, {0, 0, 0, 1, 0, 0}};
        ec = pair[axis][comp];
      }
      PwFace f{};
      f.comp = comp; f.axis = axis; f.nf = nf; f.off = off; f.sign = sign;
      f.wave = (int)q;
      for (int a = 0; a < 3; ++a) { f.lo[a] = rlo[ec][a]; f.hi[a] = rhi[ec][a]; }
      f.fixed = high ? fhi[ec][axis] : flo[ec][axis] - (comp >= 3 ? 1 : 0);
      const float *I = h->idh_[half ? axis : 3 + axis].data();
      f.id = I[f.fixed + 1];
      h->pw_faces[half].push_back(f);
    };
    // AdvancePlaneWaveE (:933-1170)
    if (il[0]) { face(0, 2, 0, false, 4, -1, -1.f); face(0, 1, 0, false, 5, -1, 1.f); }
    if (il[1]) { face(0, 2, 0, true, 4, 0, 1.f); face(0, 1, 0, true, 5, 0, -1.f); }
    if (il[2]) { face(0, 0, 1, false, 5, -1, -1.f); face(0, 2, 1, false, 3, -1, 1.f); }
    if (il[3]) { face(0, 2, 1, true, 3, 0, -1.f); face(0, 0, 1, true, 5, 0, 1.f); }
    if (il[4]) { face(0, 0, 2, false, 4, -1, 1.f); face(0, 1, 2, false, 3, -1, -1.f); }
    if (il[5]) { face(0, 0, 2, true, 4, 0, -1.f); face(0, 1, 2, true, 3, 0, 1.f); }
    // AdvancePlaneWaveH (:1213-1470)
    if (il[0]) { face(1, 5, 0, false, 1, 1, 1.f); face(1, 4, 0, false, 2, 1, -1.f); }
    if (il[1]) { face(1, 5, 0, true, 1, 0, -1.f); face(1, 4, 0, true, 2, 0, 1.f); }
    if (il[2]) { face(1, 3, 1, false, 2, 1, 1.f); face(1, 5, 1, false, 0, 1, -1.f); }
    if (il[3]) { face(1, 3, 1, true, 2, 0, -1.f); face(1, 5, 1, true, 0, 0, 1.f); }
    if (il[4]) { face(1, 3, 2, false, 1, 1, -1.f); face(1, 4, 2, false, 0, 1, 1.f); }
    if (il[5]) { face(1, 3, 2, true, 1, 0, 1.f); face(1, 4, 2, true, 0, 0, -1.f); }
  }
  int rc;
  for (size_t q = 0; q < waves.size(); ++q) {
    float *de = nullptr;
    if ((rc = upload(h, &de, evols[q].data(), evols[q].size(), &h->pw_allocs))) return rc;
    waves[q].evol = de;
  }
  if ((rc = upload(h, &h->dwaves, waves.data(), waves.size(), &h->pw_allocs))) return rc;
  if ((rc = upload(h, &h->dphys, phys.data(), phys.size(), &h->pw_allocs))) return rc;
  if ((rc = upload(h, &h->dtiempo, h->tiempo.data(), h->tiempo.size(), &h->pw_allocs))) return rc;
  h->pstride = S;
  // launch groups: per half, per wave, per face axis (faces of a group are
  // parallel planes, so no two touch the same cell); the group order is the
  // reference's face order
  h->pw_groups.clear();
  for (int half = 0; half < 2; ++half)
    for (size_t q = 0; q < h->pwin.size(); ++q)
      for (int axis = 0; axis < 3; ++axis) {
        PwGroup grp;
        grp.half = half;
        grp.start.push_back(0);
        for (const PwFace &f : h->pw_faces[half]) {
          if (f.wave != (int)q || f.axis != axis) continue;
          const int ua = f.axis == 0 ? 1 : 0, va = f.axis == 2 ? 1 : 2;
          const long long nu = f.hi[ua] - f.lo[ua] + 1, nv = f.hi[va] - f.lo[va] + 1;
          if (nu <= 0 || nv <= 0) continue;
          PwRun r{};
          r.f = f;
          r.F = h->fld[f.comp];
          r.fp16 = h->g.fp16;
          r.fs = h->g.fs[f.comp];
          r.fi = h->g.fi[f.comp];
          r.px = h->g.px;
          r.xoff = h->g.xoff;
          r.pxy = h->g.pxy;
          r.wave = h->dwaves;
          r.phys = h->dphys;
          r.pstride = S;
          r.half = half;
          r.tiempo = h->dtiempo;
          r.dn = nullptr;
          r.dt = h->dt;
          r.G = half ? h->gm2h[1] : h->g2h[1];
          r.nu = (int)nu;
          grp.runs.push_back(r);
          grp.start.push_back(grp.start.back() + nu * nv);
        }
        if (grp.runs.empty()) continue;
        std::vector<PwRun> counted = grp.runs;
        for (PwRun &r : counted) r.dn = h->dstep;
        if ((rc = upload(h, &grp.druns[0], grp.runs.data(), grp.runs.size(), &h->pw_allocs)))
          return rc;
        if ((rc = upload(h, &grp.druns[1], counted.data(), counted.size(), &h->pw_allocs)))
          return rc;
        if ((rc = upload(h, &grp.dstart, grp.start.data(), grp.start.size(), &h->pw_allocs)))
          return rc;
        h->pw_groups.push_back(std::move(grp));
      }
  h->pw_ready = true;
  return FDTD_OK;
}

// ---- frequency-domain observation ------------------------------------------
// Phase factors of observation.F90:3532-3533 per step and frequency:
// E: dt * exp(mcPI2 f t_n); H: that times exp(mcPI2 f dt / 2).  mcPI2 =
// -j 2 pi with pi in real(4) (fdetypes.F90:187-190); f is real(4), the
// products run in real(8), exp of the purely imaginary argument is
// (cos y, sin y).
void dft_phases(const DftIn &d, const std::vector<double> &tiempo, double dt,
                std::vector<double2> &aux) {
  const float pi4 = 3.14159265358979323846f;
  const double mc = -(double)(2.0f * pi4);
  const size_t nf = d.freq.size();
  aux.resize(tiempo.size() * nf);
  for (size_t n = 0; n < tiempo.size(); ++n)
    for (size_t f = 0; f < nf; ++f) {
      const double F = (double)d.freq[f];
      const double y = (mc * F) * tiempo[n];
      double re = dt * std::cos(y), im = dt * std::sin(y);
      if (d.kind == 1) {
        const double y2 = ((mc * F) * dt) * 0.5;
        const double c = std::cos(y2), sn = std::sin(y2);
        const double r2 = re * c - im * sn, i2 = re * sn + im * c;
        re = r2;
        im = i2;
      }
      aux[n * nf + f] = make_double2(re, im);
    }
}

int build_dfts(fdtd_handle *h) {
  if (h->dft_ready) return FDTD_OK;
  for (void *p : h->dft_allocs) hipFree(p);
  h->dft_allocs.clear();
  if (h->dftin.empty()) { h->dft_ready = true; return FDTD_OK; }
  if (h->tiempo.empty()) return fail(h, FDTD_ERR_STATE, "fdtd_set_time not called");
  int rc;
  for (size_t q = 0; q < h->dftin.size(); ++q) {
    std::vector<double2> aux;
    dft_phases(h->dftin[q], h->tiempo, h->dt, aux);
    double2 *da = nullptr;
    if ((rc = upload(h, &da, aux.data(), aux.size(), &h->dft_allocs))) return rc;
    h->dfts[q].aux = da;
  }
  h->dft_ready = true;
  return FDTD_OK;
}

int prepare(fdtd_handle *h) {
  int rc = check_ready(h);
  if (rc) return rc;
  if ((rc = build_ade(h))) return rc;
  if ((rc = build_mur(h))) return rc;
  if ((rc = build_nodal(h))) return rc;
  if ((rc = build_planewaves(h))) return rc;
  if ((rc = build_dfts(h))) return rc;
  return FDTD_OK;
}

int run_e(fdtd_handle *h, int yee, int cpml, int ade) {
  if (h->g.fp16 && (yee != cpml || h->g.nmedia + 1 > 4096))
    return fail(h, FDTD_ERR_STATE, "binary16 storage runs the fused passes only (<= 4095 media)");
  // the fused pass (yee and cpml) always runs the CPML kernel: without slab
  // layers it updates no psi
  const int c = cpml && (h->g.has_cpml || yee);
  const int a = ade && h->g.ade[0].has;
  const int *tmed = h->tmed_ready ? h->tmed : nullptr;
  int e = h->mode == FDTD_MODE_EXACT ? exact::launch_e(h->g, h->tm, h->tm4, tmed, yee, c, a, h->stream)
                                     : fast::launch_e(h->g, h->tm, h->tm4, tmed, yee, c, a, h->stream);
  if (e) return fail(h, FDTD_ERR_HIP, "E pass launch failed: %s", hipGetErrorString((hipError_t)e));
  return FDTD_OK;
}

int run_h(fdtd_handle *h, int yee, int cpml, int ade) {
  if (h->g.fp16 && (yee != cpml || h->g.nmedia + 1 > 4096))
    return fail(h, FDTD_ERR_STATE, "binary16 storage runs the fused passes only (<= 4095 media)");
  const int c = cpml && (h->g.has_cpml || yee);
  const int a = ade && h->g.ade[1].has;
  const int *tmed = h->tmed_ready ? h->tmed : nullptr;
  int e = h->mode == FDTD_MODE_EXACT ? exact::launch_h(h->g, h->tm, h->tm4, tmed, yee, c, a, h->stream)
                                     : fast::launch_h(h->g, h->tm, h->tm4, tmed, yee, c, a, h->stream);
  if (e) return fail(h, FDTD_ERR_HIP, "H pass launch failed: %s", hipGetErrorString((hipError_t)e));
  return FDTD_OK;
}

int run_ade(fdtd_handle *h, int side) { return side ? run_h(h, 0, 0, 1) : run_e(h, 0, 0, 1); }

// UpdateObservation's frequency-domain sums at step n (or *dn)
int run_dft(fdtd_handle *h, int n, const int *dn) {
  if (h->dfts.empty()) return FDTD_OK;
  if (!dn && (n < 0 || n >= (int)h->tiempo.size()))
    return fail(h, FDTD_ERR_ARG, "step %d outside the time vector (0..%d)", n,
                (int)h->tiempo.size() - 1);
  for (const DftDev &d : h->dfts) {
    int e = launch_dft(h->g, d, n, dn, h->stream);
    if (e) return fail(h, FDTD_ERR_HIP, "DFT launch failed: %s", hipGetErrorString((hipError_t)e));
  }
  return FDTD_OK;
}

// AdvancePlaneWaveE (half 0, after the pole stage) / AdvancePlaneWaveH
// (half 1, after the CPML), one launch per face in the reference's order
int run_pw(fdtd_handle *h, int half, int n, const int *dn) {
  if (h->pw_faces[half].empty()) return FDTD_OK;
  if (!dn && (n < 0 || n >= (int)h->tiempo.size()))
    return fail(h, FDTD_ERR_ARG, "step %d outside the time vector (0..%d)", n,
                (int)h->tiempo.size() - 1);
  for (PwGroup &grp : h->pw_groups) {
    if (grp.half != half) continue;
    // the captured step reads the device step counter, eager calls pass n
    int e = launch_tfsf_list(grp.druns[dn ? 1 : 0], grp.dstart, (int)grp.runs.size(), grp.start.back(), n,
                             h->stream);
    if (e) return fail(h, FDTD_ERR_HIP, "plane-wave launch failed: %s", hipGetErrorString((hipError_t)e));
  }
  return FDTD_OK;
}

int run_nodal(fdtd_handle *h, int half, int n, const int *dn) {
  if (h->ncells[half] == 0) return FDTD_OK;
  if (!dn && (n < 0 || n >= h->nvalues_steps))
    return fail(h, FDTD_ERR_ARG, "step %d outside the time vector (0..%d)", n, h->nvalues_steps - 1);
  int e = launch_nodal(h->g, h->dcells[half], h->ncells[half], h->dvalues, (int)h->srcs.size(),
                       n, dn, h->serial[half], h->stream);
  if (e) return fail(h, FDTD_ERR_HIP, "nodal launch failed");
  return FDTD_OK;
}

}  // namespace

extern "C" {

int fdtd_abi_version(void) { return FDTD_ABI_VERSION; }

const char *fdtd_last_error(const fdtd_handle *h) {
  return h ? h->err.c_str() : g_create_error.c_str();
}

int fdtd_create(const fdtd_grid_desc *d, fdtd_handle **out) {
  if (!d || !out) return fail(nullptr, FDTD_ERR_ARG, "null argument");
  *out = nullptr;
  if (d->nx < 1 || d->ny < 1 || d->nz < 1)
    return fail(nullptr, FDTD_ERR_ARG, "grid %dx%dx%d must be at least 1 cell per axis", d->nx, d->ny, d->nz);
  if (d->mode != FDTD_MODE_FAST && d->mode != FDTD_MODE_EXACT)
    return fail(nullptr, FDTD_ERR_ARG, "unknown mode %d", d->mode);
  if (d->storage != FDTD_STORE_FP32 && d->storage != FDTD_STORE_FP16)
    return fail(nullptr, FDTD_ERR_ARG, "unknown storage %d", d->storage);
  if (d->storage == FDTD_STORE_FP16 && d->mode != FDTD_MODE_FAST)
    return fail(nullptr, FDTD_ERR_ARG, "binary16 field storage runs in FAST mode only");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(nullptr, FDTD_ERR_DEVICE, "no HIP device available");
  if (d->device < 0 || d->device >= ndev)
    return fail(nullptr, FDTD_ERR_DEVICE, "device %d out of range (%d devices)", d->device, ndev);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, d->device) != hipSuccess)
    return fail(nullptr, FDTD_ERR_DEVICE, "cannot query device %d", d->device);
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(nullptr, FDTD_ERR_DEVICE, "device %d is %s; this library is built for gfx950",
                d->device, prop.gcnArchName);
  fdtd_handle *h = new fdtd_handle();
  h->device = d->device;
  h->mode = d->mode;
  h->storage = d->storage;
  h->lds_max = (int)std::min<size_t>(prop.sharedMemPerBlock, 1u << 30);
  if (hipSetDevice(d->device) != hipSuccess) { delete h; return fail(nullptr, FDTD_ERR_DEVICE, "hipSetDevice failed"); }
  if (d->stream) {
    h->stream = (hipStream_t)d->stream;
  } else {
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
      delete h;
      return fail(nullptr, FDTD_ERR_HIP, "hipStreamCreate failed");
    }
    h->own_stream = true;
  }
  DevGrid &g = h->g;
  g.nx = d->nx; g.ny = d->ny; g.nz = d->nz;
  // rows of grids at least 256 cells wide start on a 128-byte line; narrower
  // rows are packed to 16 bytes, so a row of a few tens of cells does not
  // fetch whole lines of padding at both ends
  if (d->nx >= 256) {
    g.xoff = kXOff;
    g.px = ((long long)kXOff + d->nx + 8 + 31) / 32 * 32;
  } else {
    g.xoff = kXOffNarrow;
    g.px = ((long long)kXOffNarrow + d->nx + 8 + 3) / 4 * 4;
  }
  g.pxy = g.px * (long long)(d->ny + 3);
  const size_t ne = field_elems(h);
  // binary16 storage: E scaled by 1 (V/m), H by 1/eta0 so both sit near 1
  g.fp16 = d->storage == FDTD_STORE_FP16 ? 1 : 0;
  for (int c = 0; c < 6; ++c) {
    g.fs[c] = (g.fp16 && c >= 3) ? 1.0f / 376.730313f : 1.0f;
    g.fi[c] = 1.0f / g.fs[c];
  }
  const size_t fes = field_bytes(h);
  int rc;
  for (int c = 0; c < 6; ++c) {
    if ((rc = dalloc(h, (void **)&h->fld[c], ne * fes)) ||
        (rc = dalloc(h, (void **)&h->mi[c], ne * 2)) || (rc = dzero(h, h->fld[c], ne * fes)) ||
        (rc = dzero(h, h->mi[c], ne * 2))) {
      g_create_error = h->err;
      fdtd_destroy(h);
      return rc;
    }
    g.f[c] = h->fld[c];
    g.mi[c] = h->mi[c];
  }
  if ((rc = dalloc(h, (void **)&h->dstep, 16)) || (rc = dzero(h, h->dstep, 16))) {
    g_create_error = h->err;
    fdtd_destroy(h);
    return rc;
  }
  // tiles of 64 x rows threads, one x cell per lane; short z chunks give
  // small grids enough blocks (64^3: 22 us per step with 4-plane chunks
  // against 41 us with 16 and 70 us with 32; 256^3: 8 planes best)
  const int planes = g.nz + 1;
  if (planes <= 160) set_tiles(h, 2, 4, 1);
  else if (planes <= 640) set_tiles(h, 4, 8, 1);
  else set_tiles(h, 4, 16, 1);
  hipStreamSynchronize(h->stream);
  *out = h;
  return FDTD_OK;
}

int fdtd_destroy(fdtd_handle *h) {
  if (!h) return FDTD_OK;
  if (int rc = refuse_batched(h)) return rc;
  hipSetDevice(h->device);
  if (h->stream) hipStreamSynchronize(h->stream);
  invalidate_graph(h);
  for (void *p : h->ade_allocs) hipFree(p);
  for (void *p : h->nodal_allocs) hipFree(p);
  for (void *p : h->mur_allocs) hipFree(p);
  for (void *p : h->step_allocs) hipFree(p);
  for (void *p : h->step_small) hipFree(p);
  for (void *p : h->pw_allocs) hipFree(p);
  for (void *p : h->dft_allocs) hipFree(p);
  for (void *p : h->dft_sums) hipFree(p);
  for (void *p : h->allocs) hipFree(p);
  if (h->stage) hipFree(h->stage);
  if (h->tmed) hipFree(h->tmed);
  if (h->solo_stream) h->stream = h->solo_stream;
  if (h->own_stream && h->stream) hipStreamDestroy(h->stream);
  delete h;
  return FDTD_OK;
}

int fdtd_set_media(fdtd_handle *h, int32_t nmedia, const float *g1, const float *g2,
                   const float *gm1, const float *gm2, const int32_t *kind) {
  if (!h || nmedia < 0 || nmedia > 32767 || !g1 || !g2 || !gm1 || !gm2 || !kind)
    return fail(h, FDTD_ERR_ARG, "bad media tables");
  if (h->media_max > nmedia)
    return fail(h, FDTD_ERR_ARG, "media index %d already in use, %d media given", h->media_max, nmedia);
  hipSetDevice(h->device);
  if (h->nmedia != nmedia) {
    for (int t = 0; t < 4; ++t) h->dg[t] = nullptr;  // old buffers stay in allocs until destroy
  }
  const float *src[4] = {g1, g2, gm1, gm2};
  int rc;
  for (int t = 0; t < 4; ++t)
    if ((rc = upload(h, &h->dg[t], src[t], (size_t)nmedia + 1))) return rc;
  h->g.g1 = h->dg[0]; h->g.g2 = h->dg[1]; h->g.gm1 = h->dg[2]; h->g.gm2 = h->dg[3];
  h->g2h.assign(g2, g2 + nmedia + 1);
  h->gm2h.assign(gm2, gm2 + nmedia + 1);
  h->kindh.assign(kind, kind + nmedia + 1);
  h->nmedia = nmedia;
  h->g.nmedia = nmedia;
  h->ade_ready = false;
  h->mur_ready = false;
  h->nodal_ready = false;
  h->pw_ready = false;
  invalidate_graph(h);
  return FDTD_OK;
}

int fdtd_set_media_index(fdtd_handle *h, int32_t comp, const int16_t *mi) {
  if (!h || comp < 0 || comp > 5 || !mi) return fail(h, FDTD_ERR_ARG, "bad media index arguments");
  hipSetDevice(h->device);
  int16_t mx = 0;
  long long ext[3];
  alloc_extent(h, comp, ext);
  const long long tot = ext[0] * ext[1] * ext[2];
  for (long long t = 0; t < tot; ++t) {
    if (mi[t] < 0) return fail(h, FDTD_ERR_ARG, "negative media index at %lld", t);
    mx = std::max(mx, mi[t]);
  }
  if (h->nmedia >= 0 && mx > h->nmedia)
    return fail(h, FDTD_ERR_ARG, "media index %d outside 0..%d", mx, h->nmedia);
  h->media_max = std::max(h->media_max, (int)mx);
  int rc = copy_box(h, comp, h->mi[comp], mi, nullptr, 2);
  if (rc) return rc;
  h->ade_ready = false;
  h->nodal_ready = false;
  invalidate_graph(h);
  return FDTD_OK;
}

int fdtd_set_spacing(fdtd_handle *h, const float *idxe, const float *idye, const float *idze,
                     const float *idxh, const float *idyh, const float *idzh) {
  if (!h || !idxe || !idye || !idze || !idxh || !idyh || !idzh)
    return fail(h, FDTD_ERR_ARG, "null spacing table");
  hipSetDevice(h->device);
  const float *src[6] = {idxe, idye, idze, idxh, idyh, idzh};
  int rc;
  for (int c = 0; c < 6; ++c) {
    // primal tables span Alloc(iHx..iHz) = n+3 entries, dual ones Alloc(iEx..
    // iEz) = n+2 (timestepping.F90:314-327); the device copy keeps n+3
    const int n = axis_n(h, c % 3) + 3;
    const int len = c < 3 ? n : n - 1;
    std::vector<float> t(src[c], src[c] + len);
    t.resize(n, 0.0f);
    if ((rc = upload(h, &h->did[c], t.data(), (size_t)n))) return rc;
    h->g.id[c] = h->did[c];
    h->idh_[c] = t;
  }
  h->nodal_ready = false;
  h->pw_ready = false;
  invalidate_graph(h);
  return FDTD_OK;
}

int fdtd_set_cpml(fdtd_handle *h, const int32_t layers[6], const float *pbe_x,
                  const float *pce_x, const float *pbm_x, const float *pcm_x,
                  const float *pbe_y, const float *pce_y, const float *pbm_y,
                  const float *pcm_y, const float *pbe_z, const float *pce_z,
                  const float *pbm_z, const float *pcm_z) {
  if (!h || !layers) return fail(h, FDTD_ERR_ARG, "null CPML arguments");
  hipSetDevice(h->device);
  DevGrid &g = h->g;
  const float *prof[12] = {pbe_x, pce_x, pbm_x, pcm_x, pbe_y, pce_y,
                           pbm_y, pcm_y, pbe_z, pce_z, pbm_z, pcm_z};
  bool any = false;
  for (int a = 0; a < 3; ++a) {
    const int n = axis_n(h, a);
    if (layers[2 * a] < 0 || layers[2 * a + 1] < 0 || layers[2 * a] + layers[2 * a + 1] > n)
      return fail(h, FDTD_ERR_ARG, "CPML layers on axis %d exceed the grid", a);
    g.L[a][0] = layers[2 * a];
    g.L[a][1] = layers[2 * a + 1];
    any = any || layers[2 * a] || layers[2 * a + 1];
    int rc;
    for (int t = 0; t < 4; ++t) {
      if (!prof[4 * a + t]) return fail(h, FDTD_ERR_ARG, "null CPML profile");
      if ((rc = upload(h, &h->dprof[4 * a + t], prof[4 * a + t], (size_t)n + 3))) return rc;
    }
    g.pbe[a] = h->dprof[4 * a]; g.pce[a] = h->dprof[4 * a + 1];
    g.pbm[a] = h->dprof[4 * a + 2]; g.pcm[a] = h->dprof[4 * a + 3];
  }
  // psi slab arrays (component sweep on the two other axes, Llo+Lhi along
  // the stretched one); slot order matches PSI_*
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int Lx = g.L[0][0] + g.L[0][1], Ly = g.L[1][0] + g.L[1][1], Lz = g.L[2][0] + g.L[2][1];
  const long long len[12] = {
      (long long)nx * Ly * (nz + 1),        // Exy
      (long long)nx * (ny + 1) * Lz,        // Exz
      (long long)Lx * ny * (nz + 1),        // Eyx
      (long long)(nx + 1) * ny * Lz,        // Eyz
      (long long)Lx * (ny + 1) * nz,        // Ezx
      (long long)(nx + 1) * Ly * nz,        // Ezy
      (long long)(nx + 1) * Ly * nz,        // Hxy
      (long long)(nx + 1) * ny * Lz,        // Hxz
      (long long)Lx * (ny + 1) * nz,        // Hyx
      (long long)nx * (ny + 1) * Lz,        // Hyz
      (long long)Lx * ny * (nz + 1),        // Hzx
      (long long)nx * Ly * (nz + 1)};       // Hzy
  for (int s = 0; s < 12; ++s) {
    int rc;
    if (len[s] != h->psilen[s]) g.psi[s] = nullptr;  // a different layout: fresh array
    if (!g.psi[s] && len[s] > 0) {
      if ((rc = dalloc(h, (void **)&g.psi[s], (size_t)len[s] * 4))) return rc;
    }
    h->psilen[s] = len[s];
    if (len[s] > 0 && (rc = dzero(h, g.psi[s], (size_t)len[s] * 4))) return rc;
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  g.has_cpml = any ? 1 : 0;
  h->pw_ready = false;
  invalidate_graph(h);
  return FDTD_OK;
}

static int set_dispersive(fdtd_handle *h, int side, int32_t ndisp, const int32_t *medium,
                          const int32_t *npoles, const double *kappa, const double *beta,
                          const double *g3) {
  if (!h || ndisp < 0 || (ndisp > 0 && (!medium || !npoles || !kappa || !beta || !g3)))
    return fail(h, FDTD_ERR_ARG, "bad dispersive arguments");
  for (int d = 0; d < ndisp; ++d)
    if (npoles[d] < 0 || npoles[d] > kPMax)
      return fail(h, FDTD_ERR_ARG, "medium %d has %d poles (max %d)", medium[d], npoles[d], kPMax);
  sfd_host::AdeIn &in = h->adein[side];
  in.ndisp = ndisp;
  in.dmed.assign(medium, medium + ndisp);
  in.dnp.assign(npoles, npoles + ndisp);
  in.dk.assign(kappa, kappa + (size_t)ndisp * kPMax * 2);
  in.db.assign(beta, beta + (size_t)ndisp * kPMax * 2);
  in.dg3.assign(g3, g3 + (size_t)ndisp * kPMax * 2);
  h->ade_ready = false;
  invalidate_graph(h);
  return FDTD_OK;
}

int fdtd_set_edispersive(fdtd_handle *h, int32_t ndisp, const int32_t *medium,
                         const int32_t *npoles, const double *kappa, const double *beta,
                         const double *g3) {
  return set_dispersive(h, 0, ndisp, medium, npoles, kappa, beta, g3);
}

int fdtd_set_mdispersive(fdtd_handle *h, int32_t ndisp, const int32_t *medium,
                         const int32_t *npoles, const double *kappa, const double *beta,
                         const double *gm3) {
  return set_dispersive(h, 1, ndisp, medium, npoles, kappa, beta, gm3);
}

int fdtd_set_mur(fdtd_handle *h, const int32_t faces[6], const float *cab1) {
  if (!h || !faces) return fail(h, FDTD_ERR_ARG, "bad Mur arguments");
  if (h->nmedia < 0) return fail(h, FDTD_ERR_STATE, "fdtd_set_media not called");
  bool any = false;
  for (int f = 0; f < 6; ++f) {
    h->mur_faces[f] = faces[f] ? 1 : 0;
    any = any || faces[f];
  }
  if (any && !cab1) return fail(h, FDTD_ERR_ARG, "Mur coefficients missing");
  if (any) h->cab1h.assign(cab1, cab1 + (size_t)6 * (h->nmedia + 1));
  else h->cab1h.clear();
  h->mur_ready = false;
  invalidate_graph(h);
  return FDTD_OK;
}

int fdtd_set_time(fdtd_handle *h, double dt, int32_t nlast, const double *tiempo) {
  if (!h || !(dt > 0.0) || nlast < 0) return fail(h, FDTD_ERR_ARG, "bad time arguments");
  h->dt = dt;
  h->g.dt = dt;
  h->tiempo.assign((size_t)nlast + 1, 0.0);
  if (tiempo) {
    std::copy(tiempo, tiempo + nlast + 1, h->tiempo.begin());
  } else {
    for (int32_t i = 1; i <= nlast; ++i) h->tiempo[i] = h->tiempo[i - 1] + dt;
  }
  h->ade_ready = false;  // FAST coefficients pre-divide by dt
  h->nodal_ready = false;
  h->pw_ready = false;
  h->dft_ready = false;
  invalidate_graph(h);
  return FDTD_OK;
}

int fdtd_add_nodal(fdtd_handle *h, int32_t comp, int32_t hard, const int32_t box[6], float amp,
                   float deltaevol, int32_t numus, const float *evol) {
  if (!h || comp < 0 || comp > 5 || !box || numus < 0 || !evol || !(deltaevol > 0.0f))
    return fail(h, FDTD_ERR_ARG, "bad nodal source");
  const int n[3] = {h->g.nx, h->g.ny, h->g.nz};
  for (int a = 0; a < 3; ++a) {
    const bool node = comp < 3 ? (comp != a) : (comp - 3 == a);
    const int hi = node ? n[a] : n[a] - 1;
    if (box[2 * a] < 0 || box[2 * a + 1] > hi || box[2 * a] > box[2 * a + 1])
      return fail(h, FDTD_ERR_ARG, "nodal box outside the sweep of component %d", comp);
  }
  Source s;
  s.comp = comp; s.hard = hard ? 1 : 0;
  std::copy(box, box + 6, s.box);
  s.amp = amp; s.deltaevol = deltaevol; s.numus = numus;
  s.evol.assign(evol, evol + numus + 1);
  h->srcs.push_back(std::move(s));
  h->nodal_ready = false;
  invalidate_graph(h);
  return FDTD_OK;
}

int fdtd_set_tiling(fdtd_handle *h, int32_t rows, int32_t zchunk, int32_t xcells) {
  if (!h || !(rows == 1 || rows == 2 || rows == 4) || zchunk < 1 || !(xcells == 1 || xcells == 4))
    return fail(h, FDTD_ERR_ARG, "tiling %d rows x %d planes x %d cells not supported", rows,
                zchunk, xcells);
  const long long nblk = (long long)((h->g.nx + kWave) / kWave) * ((h->g.ny + rows) / rows) *
                         ((h->g.nz + zchunk) / zchunk);
  if (nblk > 0x7fffffffLL) return fail(h, FDTD_ERR_ARG, "tiling gives too many blocks");
  set_tiles(h, rows, zchunk, xcells);
  invalidate_graph(h);
  return FDTD_OK;
}

int fdtd_get_tiling(const fdtd_handle *h, int32_t *rows, int32_t *zchunk, int32_t *xcells) {
  if (!h || !rows || !zchunk || !xcells) return FDTD_ERR_ARG;
  *rows = h->tm.bdy;
  *zchunk = h->tile_zchunk;
  *xcells = h->tm4.vec;
  return FDTD_OK;
}

int fdtd_time_halves(fdtd_handle *h, int32_t reps, float ms[3]) {
  if (!h || reps < 1 || !ms) return FDTD_ERR_ARG;
  if (int rb = refuse_batched(h)) return rb;
  hipSetDevice(h->device);
  int rc = prepare(h);
  if (!rc) rc = build_tile_media(h);
  if (rc) return rc;
  hipEvent_t ev[4];
  for (int i = 0; i < 4; ++i)
    if (hipEventCreate(&ev[i]) != hipSuccess) {
      for (int j = 0; j < i; ++j) hipEventDestroy(ev[j]);
      return fail(h, FDTD_ERR_HIP, "hipEventCreate failed");
    }
  // one untimed E+H step, then reps of each half
  if (!(rc = run_e(h, 1, 1, 1)) && !(rc = run_h(h, 1, 1, 1))) {
    hipEventRecord(ev[0], h->stream);
    for (int r = 0; r < reps && !rc; ++r) rc = run_e(h, 1, 1, 1);
    hipEventRecord(ev[1], h->stream);
    for (int r = 0; r < reps && !rc; ++r) rc = run_h(h, 1, 1, 1);
    hipEventRecord(ev[2], h->stream);
    for (int r = 0; r < reps && !rc; ++r) rc = run_e(h, 0, 0, 1);
    hipEventRecord(ev[3], h->stream);
  }
  if (!rc) {
    if (hipEventSynchronize(ev[3]) != hipSuccess) {
      rc = fail(h, FDTD_ERR_HIP, "timing events failed");
    } else {
      for (int i = 0; i < 3; ++i) {
        float t = 0.f;
        hipEventElapsedTime(&t, ev[i], ev[i + 1]);
        ms[i] = t / reps;
      }
    }
  }
  for (int i = 0; i < 4; ++i) hipEventDestroy(ev[i]);
  return rc;
}

int fdtd_add_dft(fdtd_handle *h, int32_t kind, const int32_t box[6], const int32_t stride[3],
                 int32_t nfreq, const float *freq, int32_t *id) {
  if (!h || (kind != 0 && kind != 1) || !box || !stride || nfreq < 1 || !freq || !id)
    return fail(h, FDTD_ERR_ARG, "bad frequency-domain probe arguments");
  hipSetDevice(h->device);
  DftIn d;
  d.kind = kind;
  DftDev dv{};
  dv.kind = kind;
  dv.nfreq = nfreq;
  long long npt = 1;
  for (int a = 0; a < 3; ++a) {
    const int lo = box[2 * a], hi = box[2 * a + 1], st = stride[a];
    if (st < 1 || lo > hi || lo < -1 || hi > axis_n(h, a) + 1)
      return fail(h, FDTD_ERR_ARG, "frequency-domain box outside the grid on axis %d", a);
    int first = lo;
    while (first <= hi && first % st != 0) ++first;
    int cnt = 0;
    for (int v = first; v <= hi; v += st) ++cnt;
    dv.first[a] = first;
    dv.stride[a] = st;
    dv.cnt[a] = cnt;
    npt *= cnt;
    d.box[2 * a] = lo; d.box[2 * a + 1] = hi; d.stride[a] = st;
  }
  if (npt == 0) return fail(h, FDTD_ERR_ARG, "frequency-domain box holds no point");
  d.freq.assign(freq, freq + nfreq);
  const size_t bytes = (size_t)npt * 3 * nfreq * sizeof(double2);
  int rc;
  if ((rc = dalloc(h, (void **)&dv.acc, bytes, &h->dft_sums))) return rc;
  HIPCHK(h, hipMemset(dv.acc, 0, bytes));
  h->dftin.push_back(std::move(d));
  h->dfts.push_back(dv);
  h->dft_ready = false;
  invalidate_graph(h);
  *id = (int32_t)h->dfts.size() - 1;
  return FDTD_OK;
}

int64_t fdtd_dft_size(const fdtd_handle *h, int32_t id) {
  if (!h || id < 0 || id >= (int32_t)h->dfts.size()) return -1;
  const DftDev &d = h->dfts[id];
  return (int64_t)d.cnt[0] * d.cnt[1] * d.cnt[2] * 3 * d.nfreq;
}

int fdtd_read_dft(fdtd_handle *h, int32_t id, double *out, int64_t capacity) {
  if (!h || id < 0 || id >= (int32_t)h->dfts.size() || !out)
    return fail(h, FDTD_ERR_ARG, "bad frequency-domain probe id");
  const int64_t n = fdtd_dft_size(h, id);
  if (capacity < n) return fail(h, FDTD_ERR_ARG, "output holds %lld of %lld complex values",
                                (long long)capacity, (long long)n);
  hipSetDevice(h->device);
  HIPCHK(h, hipMemcpyAsync(out, h->dfts[id].acc, (size_t)n * sizeof(double2),
                           hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FDTD_OK;
}

int fdtd_update_observation(fdtd_handle *h, int32_t n) {
  if (!h) return FDTD_ERR_ARG;
  if (int rb = refuse_batched(h)) return rb;
  hipSetDevice(h->device);
  int rc = prepare(h);
  if (rc) return rc;
  if ((rc = probe_room(h, 1))) return rc;
  if (h->nprobe && launch_probes(h->g, h->dprobes, h->nprobe, h->dpbuf, h->dpcount, h->pcap,
                                 h->stream))
    return fail(h, FDTD_ERR_HIP, "probe launch failed");
  h->pbuffered += h->nprobe ? 1 : 0;
  return run_dft(h, n, nullptr);
}

int fdtd_set_lines(fdtd_handle *h, const float *line_x, const float *line_y,
                   const float *line_z) {
  if (!h || !line_x || !line_y || !line_z) return fail(h, FDTD_ERR_ARG, "null line coordinates");
  const float *src[3] = {line_x, line_y, line_z};
  for (int a = 0; a < 3; ++a) h->line[a].assign(src[a], src[a] + axis_n(h, a) + 3);
  h->pw_ready = false;
  invalidate_graph(h);
  return FDTD_OK;
}

int fdtd_add_planewave(fdtd_handle *h, const int32_t box[6], int32_t nmodes, const float *p,
                       const float *e, const float *incert, float deltaevol, int32_t numus,
                       const float *evol, float eps0, float mu0) {
  if (!h || !box || nmodes < 1 || nmodes > kPwModes || !p || !e || !incert || numus < 1 ||
      !evol || !(deltaevol > 0.0f) || !(eps0 > 0.0f) || !(mu0 > 0.0f))
    return fail(h, FDTD_ERR_ARG, "bad plane-wave arguments");
  const int n[3] = {h->g.nx, h->g.ny, h->g.nz};
  for (int a = 0; a < 3; ++a)
    if (box[2 * a] < 0 || box[2 * a + 1] > n[a] || box[2 * a] >= box[2 * a + 1])
      return fail(h, FDTD_ERR_ARG, "plane-wave box outside the grid on axis %d", a);
  for (int m = 0; m < nmodes; ++m)
    if (p[3 * m] == 0.0f && p[3 * m + 1] == 0.0f && p[3 * m + 2] == 0.0f)
      return fail(h, FDTD_ERR_ARG, "plane-wave mode %d has no direction", m);
  PlaneWaveIn w{};
  std::copy(box, box + 6, w.esq);
  w.nmodes = nmodes;
  for (int m = 0; m < nmodes; ++m) {
    for (int a = 0; a < 3; ++a) { w.p[m][a] = p[3 * m + a]; w.e[m][a] = e[3 * m + a]; }
    w.incert[m] = incert[m];
  }
  w.deltaevol = deltaevol;
  w.evol.assign(evol, evol + numus + 1);
  w.eps0 = eps0;
  w.mu0 = mu0;
  h->pwin.push_back(std::move(w));
  h->pw_ready = false;
  invalidate_graph(h);
  return FDTD_OK;
}

int fdtd_advance_e(fdtd_handle *h, int32_t n) {
  if (!h) return FDTD_ERR_ARG;
  if (int rc = refuse_batched(h)) return rc;
  h->prev_exact = true;
  hipSetDevice(h->device);
  int rc;
  if ((rc = prepare(h)) || (rc = build_tile_media(h)) || (rc = run_e(h, 1, 1, 1)) || (rc = run_pw(h, 0, n, nullptr)) ||
      (rc = run_nodal(h, 0, n, nullptr)))
    return rc;
  return FDTD_OK;
}

int fdtd_advance_h(fdtd_handle *h, int32_t n) {
  if (!h) return FDTD_ERR_ARG;
  if (int rc = refuse_batched(h)) return rc;
  h->prev_exact = true;
  hipSetDevice(h->device);
  int rc;
  if ((rc = prepare(h)) || (rc = build_tile_media(h)) || (rc = run_h(h, 1, 1, 1)) || (rc = run_pw(h, 1, n, nullptr)) ||
      (rc = run_nodal(h, 1, n, nullptr)) || (rc = run_mur(h)))
    return rc;
  return FDTD_OK;
}

#define SFD_GRANULAR(name, body)                                               \
  int name {                                                                   \
    if (!h) return FDTD_ERR_ARG;                                               \
    if (int rb = refuse_batched(h)) return rb;                                 \
    h->prev_exact = true;                                                      \
    hipSetDevice(h->device);                                                   \
    int rc = prepare(h);                                                       \
    if (!rc) rc = build_tile_media(h);                                         \
    if (rc) return rc;                                                         \
    return body;                                                               \
  }
SFD_GRANULAR(fdtd_advance_e_yee(fdtd_handle *h), run_e(h, 1, 0, 0))
SFD_GRANULAR(fdtd_advance_e_cpml(fdtd_handle *h), run_e(h, 0, 1, 0))
SFD_GRANULAR(fdtd_advance_e_dispersive(fdtd_handle *h), run_e(h, 0, 0, 1))
SFD_GRANULAR(fdtd_advance_planewave_e(fdtd_handle *h, int32_t n), run_pw(h, 0, n, nullptr))
SFD_GRANULAR(fdtd_advance_nodal_e(fdtd_handle *h, int32_t n), run_nodal(h, 0, n, nullptr))
SFD_GRANULAR(fdtd_advance_h_yee(fdtd_handle *h), run_h(h, 1, 0, 0))
SFD_GRANULAR(fdtd_advance_h_cpml(fdtd_handle *h), run_h(h, 0, 1, 0))
SFD_GRANULAR(fdtd_advance_h_dispersive(fdtd_handle *h), run_h(h, 0, 0, 1))
SFD_GRANULAR(fdtd_advance_mur(fdtd_handle *h), run_mur(h))
SFD_GRANULAR(fdtd_advance_planewave_h(fdtd_handle *h, int32_t n), run_pw(h, 1, n, nullptr))
SFD_GRANULAR(fdtd_advance_nodal_h(fdtd_handle *h, int32_t n), run_nodal(h, 1, n, nullptr))
#undef SFD_GRANULAR

int fdtd_run(fdtd_handle *h, int32_t n0, int32_t nsteps) {
  if (!h || n0 < 0 || nsteps < 0) return fail(h, FDTD_ERR_ARG, "bad run arguments");
  if (int rb = refuse_batched(h)) return rb;
  hipSetDevice(h->device);
  int rc = prepare(h);
  if (rc) return rc;
  if (nsteps == 0) return FDTD_OK;
  // every step samples the probes once
  if ((rc = probe_room(h, nsteps))) return rc;
  if (!h->srcs.empty() && n0 + nsteps > h->nvalues_steps)
    return fail(h, FDTD_ERR_ARG, "steps %d..%d beyond the time vector", n0, n0 + nsteps - 1);
  if ((!h->pwin.empty() || !h->dftin.empty()) && (n0 < 0 || n0 + nsteps > (int)h->tiempo.size()))
    return fail(h, FDTD_ERR_ARG, "steps %d..%d beyond the time vector", n0, n0 + nsteps - 1);
  // the step counter lives on the device so one captured step replays for
  // every n
  HIPCHK(h, hipMemcpyAsync(h->dstep, &n0, 4, hipMemcpyHostToDevice, h->stream));
  // one-launch steps go in pairs (A -> B -> A); an odd last step, and every
  // step when they are off, replays the two-pass step
  const bool paired = h->step_rows > 0 && nsteps >= 2;
  const int pairs = paired ? nsteps / 2 : 0;
  const int singles = paired ? nsteps % 2 : nsteps;
  if (pairs > 0 && !h->pexec) {
    if ((rc = build_step(h))) return rc;
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int r2 = run_step(h, 0);
    if (!r2) r2 = run_step(h, 1);
    hipGraph_t graph = nullptr;
    hipError_t ec = hipStreamEndCapture(h->stream, &graph);
    if (r2) { if (graph) hipGraphDestroy(graph); return r2; }
    if (ec != hipSuccess) return fail(h, FDTD_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ec));
    h->pgraph = graph;
    HIPCHK(h, hipGraphInstantiate(&h->pexec, graph, nullptr, nullptr, 0));
  }
  if (pairs > 0 && (rc = sync_b(h))) return rc;
  int32_t s0 = 0;
  if (pairs > 0 && h->prev_exact) {
    // the first pair after fields were set from outside reads the stored
    // previous fields of the pole updates
    if ((rc = run_step(h, 0, true)) || (rc = run_step(h, 1, true))) return rc;
    s0 = 1;
  }
  if (pairs > 0) h->prev_exact = false;
  for (int32_t s = s0; s < pairs; ++s) HIPCHK(h, hipGraphLaunch(h->pexec, h->stream));
  h->pbuffered += 2LL * pairs;
  if (pairs > 0) {
    h->prev_lazy = true;
    if ((rc = settle_prev(h))) return rc;
  }
  if (singles == 0) return FDTD_OK;
  if ((rc = build_tile_media(h))) return rc;
  // one two-pass step, reading the step number from the device counter and
  // advancing it, recorded nrep times into one graph
  auto capture = [h](int nrep, hipGraph_t *gout, hipGraphExec_t *xout) -> int {
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int r2 = 0;
    for (int r = 0; r < nrep && !r2; ++r) {
      r2 = run_e(h, 1, 1, 1);
      if (!r2) r2 = run_pw(h, 0, 0, h->dstep);
      if (!r2) r2 = run_nodal(h, 0, 0, h->dstep);
      if (!r2) r2 = run_h(h, 1, 1, 1);
      if (!r2) r2 = run_pw(h, 1, 0, h->dstep);
      if (!r2) r2 = run_nodal(h, 1, 0, h->dstep);
      if (!r2) r2 = run_mur(h);
      if (!r2 && h->nprobe)
        r2 = launch_probes(h->g, h->dprobes, h->nprobe, h->dpbuf, h->dpcount, h->pcap, h->stream)
                 ? fail(h, FDTD_ERR_HIP, "probe launch failed") : 0;
      if (!r2) r2 = run_dft(h, 0, h->dstep);
      if (!r2) r2 = launch_tick(h->dstep, h->stream) ? fail(h, FDTD_ERR_HIP, "tick failed") : 0;
    }
    hipGraph_t graph = nullptr;
    hipError_t ec = hipStreamEndCapture(h->stream, &graph);
    if (r2) { if (graph) hipGraphDestroy(graph); return r2; }
    if (ec != hipSuccess) return fail(h, FDTD_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ec));
    *gout = graph;
    HIPCHK(h, hipGraphInstantiate(xout, graph, nullptr, nullptr, 0));
    return FDTD_OK;
  };
  const int K = h->graph_steps;
  int32_t left = singles;
  if (K > 1 && left >= K) {
    if (!h->mexec && (rc = capture(K, &h->mgraph, &h->mexec))) return rc;
    for (; left >= K; left -= K) HIPCHK(h, hipGraphLaunch(h->mexec, h->stream));
  }
  if (left > 0 && !h->graph_ok) {
    if (h->gexec) hipGraphExecDestroy(h->gexec);
    if (h->graph) hipGraphDestroy(h->graph);
    h->gexec = nullptr;
    h->graph = nullptr;
    if ((rc = capture(1, &h->graph, &h->gexec))) return rc;
    h->graph_ok = true;
  }
  for (int32_t s = 0; s < left; ++s) HIPCHK(h, hipGraphLaunch(h->gexec, h->stream));
  h->pbuffered += singles;
  return FDTD_OK;
}

int fdtd_set_graph_steps(fdtd_handle *h, int32_t steps) {
  if (!h) return FDTD_ERR_ARG;
  if (steps < 1 || steps > 64) return fail(h, FDTD_ERR_ARG, "graph steps %d outside 1..64", steps);
  if (int rb = refuse_batched(h)) return rb;
  if (steps != h->graph_steps) {
    hipSetDevice(h->device);
    if (h->stream) hipStreamSynchronize(h->stream);
    drop_multi_graph(h);
    h->graph_steps = steps;
  }
  return FDTD_OK;
}

int fdtd_set_step_mode(fdtd_handle *h, int32_t rows, int32_t zchunk) {
  if (!h || !(rows == 0 || rows == 4 || rows == 8 || rows == 16) || zchunk < 1)
    return fail(h, FDTD_ERR_ARG, "step mode %d rows x %d planes not supported", rows, zchunk);
  const int kc_before = h->step_kc;
  h->step_kc = zchunk;
  if (rows && !step_supported(h, rows)) {
    h->step_kc = kc_before;
    return fail(h, FDTD_ERR_STATE, "one-launch steps need fp32 storage, small media tables and at most 64 TF/SF faces per half");
  }
  h->step_rows = rows;
  invalidate_graph(h);
  return FDTD_OK;
}

int fdtd_set_step_class_rows(fdtd_handle *h, int32_t classes, int32_t rows) {
  if (!h || classes < 0 || classes >= (1 << FDTD_STEP_CLASSES) ||
      !(rows == 0 || rows == 4 || rows == 6 || rows == 8 || rows == 16))
    return fail(h, FDTD_ERR_ARG, "class rows %d (classes 0x%x) not supported", rows, classes);
  for (int c = 0; c < FDTD_STEP_CLASSES; ++c)
    if (classes >> c & 1) h->step_creq[c] = rows;
  invalidate_graph(h);
  return FDTD_OK;
}

int fdtd_get_step_class_rows(const fdtd_handle *h, int32_t rows[FDTD_STEP_CLASSES]) {
  if (!h || !rows) return FDTD_ERR_ARG;
  for (int c = 0; c < FDTD_STEP_CLASSES; ++c) rows[c] = class_rows(h, c);
  return FDTD_OK;
}

int fdtd_set_step_vector(fdtd_handle *h, int32_t on) {
  if (!h || (on != 0 && on != 1)) return fail(h, FDTD_ERR_ARG, "step vector flag %d", on);
  h->step_vec = on != 0;
  invalidate_graph(h);
  return FDTD_OK;
}

int fdtd_get_step_mode(const fdtd_handle *h, int32_t *rows, int32_t *zchunk) {
  if (!h || !rows || !zchunk) return FDTD_ERR_ARG;
  *rows = h->step_rows;
  *zchunk = h->step_kc;
  return FDTD_OK;
}

int fdtd_time_step(fdtd_handle *h, int32_t reps, float *ms) {
  if (!h || reps < 1 || !ms) return FDTD_ERR_ARG;
  if (int rb = refuse_batched(h)) return rb;
  if (h->step_rows == 0) return fail(h, FDTD_ERR_STATE, "one-launch steps are off");
  hipSetDevice(h->device);
  int rc;
  if ((rc = prepare(h)) || (rc = build_step(h)) || (rc = sync_b(h))) return rc;
  hipEvent_t ev[2];
  if (hipEventCreate(&ev[0]) != hipSuccess) return fail(h, FDTD_ERR_HIP, "hipEventCreate failed");
  if (hipEventCreate(&ev[1]) != hipSuccess) {
    hipEventDestroy(ev[0]);
    return fail(h, FDTD_ERR_HIP, "hipEventCreate failed");
  }
  // the step kernels alone, in pairs so that the state ends in A
  auto pair = [&]() -> int {
    for (int dir = 0; dir < 2; ++dir) {
      const int e = launch_step_kernels(h, dir);
      if (e) return e;
    }
    return FDTD_OK;
  };
  rc = pair();
  if (!rc) {
    hipEventRecord(ev[0], h->stream);
    for (int r = 0; r < reps && !rc; ++r) rc = pair();
    hipEventRecord(ev[1], h->stream);
  }
  h->prev_lazy = true;
  if (!rc) rc = settle_prev(h);
  if (!rc) {
    if (hipEventSynchronize(ev[1]) != hipSuccess) {
      rc = fail(h, FDTD_ERR_HIP, "timing events failed");
    } else {
      float t = 0.f;
      hipEventElapsedTime(&t, ev[0], ev[1]);
      *ms = t / (2.0f * reps);
    }
  }
  hipEventDestroy(ev[0]);
  hipEventDestroy(ev[1]);
  return rc;
}

int fdtd_step_classes(fdtd_handle *h, int32_t reps, int32_t cls[FDTD_STEP_CLASSES],
                      double cells[FDTD_STEP_CLASSES], double bytes[FDTD_STEP_CLASSES],
                      float ms[FDTD_STEP_CLASSES]) {
  if (!h || reps < 0 || !cls || !cells || !bytes || (reps > 0 && !ms)) return FDTD_ERR_ARG;
  if (int rb = refuse_batched(h)) return rb;
  if (h->step_rows == 0) return fail(h, FDTD_ERR_STATE, "one-launch steps are off");
  hipSetDevice(h->device);
  int rc;
  if ((rc = prepare(h)) || (rc = build_step(h)) || (rc = sync_b(h))) return rc;
  const int nc = (int)h->step_cls.size();
  for (int q = 0; q < nc; ++q) {
    cls[q] = h->step_cls[q];
    cells[q] = h->step_cells[q];
    bytes[q] = h->step_bytes[q];
  }
  if (reps == 0) return nc;
  hipEvent_t ev[2];
  if (hipEventCreate(&ev[0]) != hipSuccess) return fail(h, FDTD_ERR_HIP, "hipEventCreate failed");
  if (hipEventCreate(&ev[1]) != hipSuccess) {
    hipEventDestroy(ev[0]);
    return fail(h, FDTD_ERR_HIP, "hipEventCreate failed");
  }
  // the incident-field buffer of the faces is filled once for step n
  if (h->step_nf > 0 &&
      launch_pw_incident(h->dsfaces, h->step_nf0, h->step_nf, h->step_maxcells, h->dstepsrc,
                         h->dstep, h->dinc, h->stream))
    rc = fail(h, FDTD_ERR_HIP, "incident-field launch failed");
  for (int q = 0; q < nc && !rc; ++q) {
    // as in the step: tiles without sources leave the stored previous field
    // of B, which nothing reads before it is written again
    StepDev sd = class_sd(h, 0, (size_t)q);
    sd.prev_from_field = 1;
    sd.lazy_prev = 1;
    auto one = [&]() { return launch_class(h, sd, (size_t)q); };
    if (one()) { rc = fail(h, FDTD_ERR_HIP, "step launch failed"); break; }
    hipEventRecord(ev[0], h->stream);
    for (int r = 0; r < reps && !rc; ++r)
      if (one()) rc = fail(h, FDTD_ERR_HIP, "step launch failed");
    hipEventRecord(ev[1], h->stream);
    if (rc) break;
    if (hipEventSynchronize(ev[1]) != hipSuccess) {
      rc = fail(h, FDTD_ERR_HIP, "timing events failed");
      break;
    }
    float t = 0.f;
    hipEventElapsedTime(&t, ev[0], ev[1]);
    ms[q] = t / reps;
  }
  hipEventDestroy(ev[0]);
  hipEventDestroy(ev[1]);
  // only the second copy was written: the grid's state (A) is unchanged
  return rc ? rc : nc;
}

int fdtd_set_probes(fdtd_handle *h, int32_t nprobe, const int32_t *comp, const int32_t *i,
                    const int32_t *j, const int32_t *k, int32_t bufsteps) {
  if (!h || nprobe < 0 || bufsteps < 1 || (nprobe && (!comp || !i || !j || !k)))
    return fail(h, FDTD_ERR_ARG, "bad probe arguments");
  hipSetDevice(h->device);
  std::vector<ProbeDev> pr(nprobe);
  const int n[3] = {h->g.nx, h->g.ny, h->g.nz};
  for (int p = 0; p < nprobe; ++p) {
    const int crd[3] = {i[p], j[p], k[p]};
    if (comp[p] < 0 || comp[p] > 5) return fail(h, FDTD_ERR_ARG, "probe %d: bad component", p);
    for (int a = 0; a < 3; ++a)
      if (crd[a] < -1 || crd[a] > n[a] + 1) return fail(h, FDTD_ERR_ARG, "probe %d outside the grid", p);
    pr[p].comp = comp[p];
    pr[p].p = (long long)(k[p] + 1) * h->g.pxy + (long long)(j[p] + 1) * h->g.px + h->g.xoff + i[p];
  }
  int rc;
  h->dprobes = nullptr; h->dpbuf = nullptr; h->dpcount = nullptr;
  if ((rc = dalloc(h, (void **)&h->dprobes, sizeof(ProbeDev) * std::max(nprobe, 1)))) return rc;
  if ((rc = dalloc(h, (void **)&h->dpbuf, (size_t)bufsteps * std::max(nprobe, 1) * 4))) return rc;
  if ((rc = dalloc(h, (void **)&h->dpcount, 16))) return rc;
  if (nprobe) HIPCHK(h, hipMemcpy(h->dprobes, pr.data(), sizeof(ProbeDev) * nprobe, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemset(h->dpcount, 0, 16));
  h->nprobe = nprobe;
  h->pcap = bufsteps;
  h->pbuffered = 0;
  invalidate_graph(h);
  return FDTD_OK;
}

int fdtd_sample_probes(fdtd_handle *h) {
  if (!h) return FDTD_ERR_ARG;
  if (!h->nprobe) return FDTD_OK;
  if (int rc = refuse_batched(h)) return rc;
  if (int rc = probe_room(h, 1)) return rc;
  hipSetDevice(h->device);
  if (launch_probes(h->g, h->dprobes, h->nprobe, h->dpbuf, h->dpcount, h->pcap, h->stream))
    return fail(h, FDTD_ERR_HIP, "probe launch failed");
  h->pbuffered += 1;
  return FDTD_OK;
}

int fdtd_read_probes(fdtd_handle *h, float *out, int32_t max_samples, int32_t *nsamples) {
  if (!h || !nsamples || (max_samples > 0 && !out)) return fail(h, FDTD_ERR_ARG, "bad probe read");
  hipSetDevice(h->device);
  *nsamples = 0;
  if (!h->nprobe) return FDTD_OK;
  int cnt = 0;
  HIPCHK(h, hipMemcpyAsync(&cnt, h->dpcount, 4, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (cnt > max_samples) return fail(h, FDTD_ERR_ARG, "%d samples buffered, room for %d", cnt, max_samples);
  if (cnt) HIPCHK(h, hipMemcpy(out, h->dpbuf, (size_t)cnt * h->nprobe * 4, hipMemcpyDeviceToHost));
  const int zero = 0;
  HIPCHK(h, hipMemcpyAsync(h->dpcount, &zero, 4, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->pbuffered = 0;
  *nsamples = cnt;
  return FDTD_OK;
}

int fdtd_get_field(fdtd_handle *h, int32_t comp, float *host) {
  if (!h || comp < 0 || comp > 5 || !host) return fail(h, FDTD_ERR_ARG, "bad field read");
  hipSetDevice(h->device);
  return field_io(h, comp, nullptr, host);
}

int fdtd_set_field(fdtd_handle *h, int32_t comp, const float *host) {
  if (h) h->step_stale = h->prev_exact = true;
  if (!h || comp < 0 || comp > 5 || !host) return fail(h, FDTD_ERR_ARG, "bad field write");
  hipSetDevice(h->device);
  return field_io(h, comp, host, nullptr);
}

// ---- field read that allocates nothing -------------------------------------
// fdtd_get_field stages through a buffer it allocates (and grows) on first
// use.  fdtd_get_field_noalloc stages through `scratch`, a file-scope device
// array of kScratchBytes that every device the library runs on holds until the
// process ends: runs of z planes are gathered into it as fp32 and copied to the
// host, so the call makes no hipMalloc/hipFree.  In host code `scratch` names a
// host-side stand-in, never the device array: the device address comes from
// hipGetSymbolAddress, once per device with that device current, and only that
// address is passed to kernels and copies.  All calls on one device share its
// array; overlapping calls on one device race.
namespace {
constexpr size_t kScratchBytes = (size_t)16 << 20;
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch_ptr[kScratchDevices] = {};
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// device address of `scratch` on h's device (current), looked up on first use
static int scratch_of(fdtd_handle *h, float **out) {
  if (h->device < 0 || h->device >= kScratchDevices)
    return fail(h, FDTD_ERR_DEVICE, "device %d beyond the %d devices the scratch table holds",
                h->device, kScratchDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[h->device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(h, FDTD_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", h->device,
                  hipGetErrorString(e));
    g_scratch_ptr[h->device] = (float *)p;
  }
  *out = g_scratch_ptr[h->device];
  return FDTD_OK;
}

int fdtd_get_field_noalloc(fdtd_handle *h, int32_t comp, float *host) {
  if (!h || comp < 0 || comp > 5 || !host) return fail(h, FDTD_ERR_ARG, "bad field read");
  HIPCHK(h, hipSetDevice(h->device));
  float *stage = nullptr;
  if (int rc = scratch_of(h, &stage)) return rc;
  long long ext[3];
  alloc_extent(h, comp, ext);
  const size_t plane = (size_t)(ext[0] * ext[1]) * 4;
  if (plane > kScratchBytes)
    return fail(h, FDTD_ERR_ARG,
                "a z plane of component %d holds %zu bytes as fp32; the no-allocation read "
                "stages at most %zu bytes (use fdtd_get_field)",
                comp, plane, kScratchBytes);
  const long long per = (long long)(kScratchBytes / plane);
  const size_t es = field_bytes(h);
  const float scale = h->storage == FDTD_STORE_FP16 ? h->g.fs[comp] : 1.0f;
  for (long long k0 = 0; k0 < ext[2]; k0 += per) {
    const long long nk = std::min(per, ext[2] - k0);
    const long long sub[3] = {ext[0], ext[1], nk};
    const char *g = (const char *)h->fld[comp] + (size_t)k0 * (size_t)h->g.pxy * es;
    HIPCHK(h, (hipError_t)launch_box_gather_f32(g, es, scale, stage, sub, h->g.px, h->g.pxy, h->g.xoff,
                                                h->stream));
    HIPCHK(h, hipMemcpyAsync(host + (size_t)k0 * (plane / 4), stage, plane * (size_t)nk,
                             hipMemcpyDeviceToHost, h->stream));
    // the next run overwrites the scratch
    HIPCHK(h, hipStreamSynchronize(h->stream));
  }
  return FDTD_OK;
}

int fdtd_set_field_scale(fdtd_handle *h, int32_t comp, float scale) {
  if (!h || comp < 0 || comp > 5 || !(scale > 0.0f) || !std::isfinite(scale))
    return fail(h, FDTD_ERR_ARG, "bad field scale");
  if (h->storage != FDTD_STORE_FP16) return FDTD_OK;  // fp32 fields carry no scale
  hipSetDevice(h->device);
  // keep the stored values: read at the old scale, write at the new one
  long long ext[3];
  alloc_extent(h, comp, ext);
  std::vector<float> tmp((size_t)(ext[0] * ext[1] * ext[2]));
  int rc = field_io(h, comp, nullptr, tmp.data());
  if (rc) return rc;
  h->g.fs[comp] = scale;
  h->g.fi[comp] = 1.0f / scale;
  rc = field_io(h, comp, tmp.data(), nullptr);
  if (rc) return rc;
  h->pw_ready = false;
  invalidate_graph(h);
  return FDTD_OK;
}

int fdtd_synchronize(fdtd_handle *h) {
  if (!h) return FDTD_ERR_ARG;
  hipSetDevice(h->device);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FDTD_OK;
}

int64_t fdtd_device_bytes(const fdtd_handle *h) { return h ? h->bytes : 0; }

int fdtd_device_memory(int32_t device, int64_t *free_bytes, int64_t *total_bytes) {
  if (!free_bytes || !total_bytes) return fail(nullptr, FDTD_ERR_ARG, "null memory outputs");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
    return fail(nullptr, FDTD_ERR_DEVICE, "no HIP device %d", device);
  int cur = 0;
  hipGetDevice(&cur);
  size_t f = 0, t = 0;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipMemGetInfo(&f, &t);
  hipSetDevice(cur);
  if (e != hipSuccess) return fail(nullptr, FDTD_ERR_HIP, "hipMemGetInfo: %s", hipGetErrorString(e));
  *free_bytes = (int64_t)f;
  *total_bytes = (int64_t)t;
  return FDTD_OK;
}

int64_t fdtd_dispersive_edges(const fdtd_handle *h, int32_t comp) {
  if (!h || comp < 0 || comp > 5 || !h->ade_ready) return -1;
  const int side = comp / 3;
  return h->g.ade[side].has ? (int64_t)h->adein[side].nedges[comp % 3] : 0;
}

int fdtd_media_fill(fdtd_handle *h, int16_t medium) {
  if (!h || medium < 0) return fail(h, FDTD_ERR_ARG, "bad medium");
  h->media_max = medium;  // every previous index is overwritten
  hipSetDevice(h->device);
  for (int c = 0; c < 6; ++c)
    if (launch_fill_i16(h->mi[c], (long long)field_elems(h), medium, h->stream))
      return fail(h, FDTD_ERR_HIP, "fill launch failed");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->ade_ready = false; h->nodal_ready = false;
  invalidate_graph(h);
  return FDTD_OK;
}

int fdtd_media_pec_walls(fdtd_handle *h, int32_t faces) {
  if (!h) return FDTD_ERR_ARG;
  hipSetDevice(h->device);
  if (launch_pec_walls(h->g, h->mi, faces, h->stream)) return fail(h, FDTD_ERR_HIP, "pec launch failed");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->ade_ready = false; h->nodal_ready = false;
  invalidate_graph(h);
  return FDTD_OK;
}

static int paint(fdtd_handle *h, int shape, const float par[12], int16_t medium, int32_t nprio,
                 const int32_t *prio) {
  if (!h || !prio || medium < 0 || medium >= nprio) return fail(h, FDTD_ERR_ARG, "bad paint arguments");
  h->media_max = std::max(h->media_max, (int)medium);
  hipSetDevice(h->device);
  int32_t *dprio = nullptr;
  HIPCHK(h, hipMalloc((void **)&dprio, (size_t)nprio * 4));
  hipError_t e = hipMemcpy(dprio, prio, (size_t)nprio * 4, hipMemcpyHostToDevice);
  // media indices beyond the priority table are left untouched by the kernel
  if (e == hipSuccess && launch_paint(h->g, h->mi, shape, par, medium, dprio, nprio, h->stream))
    e = hipErrorLaunchFailure;
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  hipFree(dprio);
  if (e != hipSuccess) return fail(h, FDTD_ERR_HIP, "paint failed: %s", hipGetErrorString(e));
  h->ade_ready = false; h->nodal_ready = false;
  invalidate_graph(h);
  return FDTD_OK;
}

int fdtd_media_paint_sphere(fdtd_handle *h, float cx, float cy, float cz, float r, int16_t medium,
                            int32_t nprio, const int32_t *prio) {
  if (!h || !(r > 0.0f)) return fail(h, FDTD_ERR_ARG, "bad sphere");
  if (h->nmedia >= 0 && nprio < h->nmedia + 1) return fail(h, FDTD_ERR_ARG, "priority table too short");
  float par[12] = {cx, cy, cz, r, 0, 0};
  const int x0 = std::max(0, (int)std::floor(cx - r) - 1), y0 = std::max(0, (int)std::floor(cy - r) - 1),
            z0 = std::max(0, (int)std::floor(cz - r) - 1);
  const int x1 = std::min(h->g.nx, (int)std::ceil(cx + r) + 1), y1 = std::min(h->g.ny, (int)std::ceil(cy + r) + 1),
            z1 = std::min(h->g.nz, (int)std::ceil(cz + r) + 1);
  par[6] = (float)x0; par[7] = (float)y0; par[8] = (float)z0;
  par[9] = (float)(x1 - x0 + 1); par[10] = (float)(y1 - y0 + 1); par[11] = (float)(z1 - z0 + 1);
  return paint(h, 0, par, medium, nprio, prio);
}

int fdtd_media_paint_box(fdtd_handle *h, int32_t x0, int32_t x1, int32_t y0, int32_t y1, int32_t z0,
                         int32_t z1, int16_t medium, int32_t nprio, const int32_t *prio) {
  if (!h || x1 <= x0 || y1 <= y0 || z1 <= z0) return fail(h, FDTD_ERR_ARG, "bad box");
  if (h->nmedia >= 0 && nprio < h->nmedia + 1) return fail(h, FDTD_ERR_ARG, "priority table too short");
  float par[12] = {(float)x0, (float)x1, (float)y0, (float)y1, (float)z0, (float)z1};
  const int lx = std::max(0, x0), ly = std::max(0, y0), lz = std::max(0, z0);
  const int hx = std::min(h->g.nx, x1), hy = std::min(h->g.ny, y1), hz = std::min(h->g.nz, z1);
  par[6] = (float)lx; par[7] = (float)ly; par[8] = (float)lz;
  par[9] = (float)(hx - lx + 1); par[10] = (float)(hy - ly + 1); par[11] = (float)(hz - lz + 1);
  return paint(h, 1, par, medium, nprio, prio);
}

}  // extern "C"

// ---- checkpoint ------------------------------------------------------------
// The complete time-stepping state of a grid: the six fields (Alloc boxes, as
// StoreFields writes them, resuming.F90:394-496), the CPML psi arrays (the
// reference's StoreFieldsCPMLBorders counterpart), the pole state (previous
// field and pole currents of the electric then the magnetic media,
// StoreFieldsEDispersives, StoreFieldsMdispersives) and the frequency-domain
// sums,
// behind a header naming the step.  The layout of the pole state depends on
// the mode, so a state loads into a grid of the same set-up and mode.
namespace {

struct StateHeader {
  char magic[4];
  int32_t version;
  int32_t mode;
  int32_t nx, ny, nz;
  int32_t step;
  int32_t nsections;
};

struct Section {
  void *dev;
  size_t bytes;
  int comp;   // >= 0: a field copied through its Alloc box
};

std::vector<Section> state_sections(fdtd_handle *h) {
  std::vector<Section> v;
  for (int c = 0; c < 6; ++c) {
    long long ext[3];
    alloc_extent(h, c, ext);
    v.push_back({h->fld[c], (size_t)(ext[0] * ext[1] * ext[2]) * field_bytes(h), c});
  }
  for (int s = 0; s < 12; ++s)
    if (h->g.psi[s] && h->psilen[s] > 0) v.push_back({h->g.psi[s], (size_t)h->psilen[s] * 4, -1});
  for (int side = 0; side < 2; ++side) {
    const AdeSet &ad = h->g.ade[side];
    if (!ad.has) continue;
    for (int c = 0; c < 3; ++c) {
      const size_t ns = (size_t)std::max(ad.nstate[c], 1LL);
      v.push_back({ad.prev[c], ns * 4, -1});
      v.push_back({ad.jst[c], h->adein[side].jbytes[c], -1});
    }
  }
  for (const MurRun &r : h->mur_runs) v.push_back({r.past, (size_t)2 * r.nu * r.nv * 4, -1});
  for (const DftDev &d : h->dfts)
    v.push_back({d.acc, (size_t)d.cnt[0] * d.cnt[1] * d.cnt[2] * 3 * d.nfreq * sizeof(double2), -1});
  return v;
}

}  // namespace

extern "C" {

int64_t fdtd_state_size(fdtd_handle *h) {
  if (!h) return -1;
  hipSetDevice(h->device);
  if (prepare(h)) return -1;
  int64_t n = sizeof(StateHeader);
  for (const Section &s : state_sections(h)) n += (int64_t)s.bytes;
  return n;
}

int fdtd_save_state(fdtd_handle *h, int32_t n, void *buf, int64_t size) {
  if (!h || !buf) return fail(h, FDTD_ERR_ARG, "null state buffer");
  const int64_t need = fdtd_state_size(h);
  if (need < 0) return FDTD_ERR_STATE;
  if (size < need) return fail(h, FDTD_ERR_ARG, "state needs %lld bytes, %lld given",
                                 (long long)need, (long long)size);
  std::vector<Section> secs = state_sections(h);
  StateHeader hd{{'S', 'F', 'D', 'S'}, 1, h->mode * 2 + (h->g.fp16 ? 1 : 0), h->g.nx, h->g.ny, h->g.nz, n,
                 (int32_t)secs.size()};
  char *p = (char *)buf;
  std::memcpy(p, &hd, sizeof hd);
  p += sizeof hd;
  for (const Section &s : secs) {
    if (s.comp >= 0) {
      int rc = copy_box(h, s.comp, s.dev, nullptr, p, field_bytes(h));
      if (rc) return rc;
    } else {
      HIPCHK(h, hipMemcpyAsync(p, s.dev, s.bytes, hipMemcpyDeviceToHost, h->stream));
    }
    p += s.bytes;
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FDTD_OK;
}

int fdtd_load_state(fdtd_handle *h, const void *buf, int64_t size, int32_t *n) {
  if (h) h->step_stale = h->prev_exact = true;
  if (!h || !buf || !n) return fail(h, FDTD_ERR_ARG, "null state buffer");
  const int64_t need = fdtd_state_size(h);
  if (need < 0) return FDTD_ERR_STATE;
  StateHeader hd;
  if (size < (int64_t)sizeof hd) return fail(h, FDTD_ERR_ARG, "state too short");
  std::memcpy(&hd, buf, sizeof hd);
  std::vector<Section> secs = state_sections(h);
  if (std::memcmp(hd.magic, "SFDS", 4) != 0 || hd.version != 1)
    return fail(h, FDTD_ERR_ARG, "not a version-1 state");
  if (hd.mode != h->mode * 2 + (h->g.fp16 ? 1 : 0) || hd.nx != h->g.nx || hd.ny != h->g.ny || hd.nz != h->g.nz ||
      hd.nsections != (int32_t)secs.size() || size != need)
    return fail(h, FDTD_ERR_ARG, "state of a different grid, mode or set-up");
  const char *p = (const char *)buf + sizeof hd;
  for (const Section &s : secs) {
    if (s.comp >= 0) {
      int rc = copy_box(h, s.comp, s.dev, p, nullptr, field_bytes(h));
      if (rc) return rc;
    } else {
      HIPCHK(h, hipMemcpyAsync(s.dev, p, s.bytes, hipMemcpyHostToDevice, h->stream));
    }
    p += s.bytes;
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  *n = hd.step;
  return FDTD_OK;
}

}  // extern "C"
