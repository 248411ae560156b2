// fdtd_aux.hip -- small gfx950 kernels around the two passes: nodal sources,
// TF/SF plane-wave faces, point probes, dispersive-edge counting for the compact pole state, and the
// device-side scene builders used for grids too large to voxelise on the host.
// Built with -ffp-contract=off: the soft-source update keeps the reference's
// rounding order.
#include <hip/hip_runtime.h>
#include <algorithm>

#include "fdtd_internal.h"

namespace sfd {

__device__ __forceinline__ long long fidx(const DevGrid &g, int i, int j, int k) {
  return (long long)(k + 1) * g.pxy + (long long)(j + 1) * g.px + g.xoff + i;
}

// AdvanceNodalE/H (nodalsources.F90:396-777).  cells are in the reference's
// processing order; values[n*nsrc + s] = evolucion(tiempo) of source s at
// step n, evaluated on the host in real(4).  Hard: F = amp*f;
// soft: F = F - ((G2*Ia*Ib)*amp)*f.
__global__ void k_nodal(DevGrid g, const NodalCell *cells, int ncells,
                        const float *values, int nsrc, int n, const int *dn,
                        int serial) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (dn) n = *dn;  // graph replay: step index lives on the device
  const int lo = serial ? 0 : t, hi = serial ? (t == 0 ? ncells : 0) : min(t + 1, ncells);
  for (int c = lo; c < hi; ++c) {
    const NodalCell cell = cells[c];
    if (cell.skip) continue;
    const float v = values[(long long)n * nsrc + cell.src];
    field_st(g, cell.comp, cell.p,
             nodal_apply(cell.hard ? 0.0f : field_ld(g, cell.comp, cell.p), cell, v));
  }
}

int launch_nodal(const DevGrid &g, const NodalCell *cells, int ncells,
                 const float *values, int nsrc, int n, const int *dn, int serial,
                 void *stream) {
  if (ncells <= 0) return 0;
  const int threads = 64;
  const int blocks = serial ? 1 : (ncells + threads - 1) / threads;
  hipLaunchKernelGGL(k_nodal, dim3(blocks), dim3(threads), 0, (hipStream_t)stream, g,
                     cells, ncells, values, nsrc, n, dn, serial);
  return (int)hipGetLastError();
}

// ---- TF/SF plane waves (planewaves.F90) ----------------------------------
// A group of TF/SF faces of AdvancePlaneWaveE/H (planewaves.F90:878-1475),
// possibly of several grids, in one launch: thread t takes cell t of the
// concatenated faces (the face is found by binary search in the prefix of face
// areas), evaluates Incid (:778-853) at the neighbouring incident component and
// applies F = F +- G * Incid * Id.
__global__ void __launch_bounds__(256) k_tfsf_list(const PwRun *runs, const long long *start,
                                                   int nruns, long long total, int n) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  int lo = 0, hi = nruns - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= t) lo = mid; else hi = mid - 1;
  }
  const PwRun &r = runs[lo];
  const PwFace &f = r.f;
  const long long c = t - start[lo];
  const int ua = f.axis == 0 ? 1 : 0;
  const int va = f.axis == 2 ? 1 : 2;
  int cc[3];
  cc[f.axis] = f.fixed;
  cc[ua] = f.lo[ua] + (int)(c % r.nu);
  cc[va] = f.lo[va] + (int)(c / r.nu);
  const int step = r.dn ? *r.dn : n;
  const float tm = r.half ? (float)(r.tiempo[step] + (double)0.5f * r.dt) : (float)r.tiempo[step];
  const float ehi = pw_incid(f, r.wave, r.phys, r.pstride, tm, cc);
  const long long p = (long long)(cc[2] + 1) * r.pxy + (long long)(cc[1] + 1) * r.px + r.xoff + cc[0];
  if (r.fp16) {
    __half *F = reinterpret_cast<__half *>(r.F);
    const float old = __half2float(F[p]) * r.fs;
    F[p] = __float2half_rn(pw_apply(old, f.sign, r.G, ehi, f.id) * r.fi);
  } else {
    float *F = r.F;
    F[p] = pw_apply(F[p], f.sign, r.G, ehi, f.id);
  }
}

int launch_tfsf_list(const PwRun *runs, const long long *start, int nruns, long long total,
                     int n, void *stream) {
  if (total <= 0 || nruns <= 0) return 0;
  const unsigned blocks = (unsigned)((total + 255) / 256);
  hipLaunchKernelGGL(k_tfsf_list, dim3(blocks), dim3(256), 0, (hipStream_t)stream, runs, start,
                     nruns, total, n);
  return (int)hipGetLastError();
}

// A dense x-fastest box of ex*ey*ez elements moved between a packed buffer
// and a grid array (rows of px elements, planes of pxy, the box's first
// element at x offset xoff - 1 of row 0, plane 0): host transfers then take
// one contiguous copy instead of one pitched copy per plane.
template <class T>
__global__ void __launch_bounds__(256) k_box_move(T *grid, T *packed, long long ex, long long ey,
                                                 long long ez, long long px, long long pxy,
                                                 int xoff, int to_grid) {
  const long long n = ex * ey * ez;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const long long i = t % ex, j = (t / ex) % ey, k = t / (ex * ey);
    const long long p = k * pxy + j * px + (xoff - 1) + i;
    if (to_grid)
      grid[p] = packed[t];
    else
      packed[t] = grid[p];
  }
}

int launch_box_move(void *grid, void *packed, size_t es, const long long ext[3], long long px,
                    long long pxy, int xoff, int to_grid, void *stream) {
  const long long n = ext[0] * ext[1] * ext[2];
  if (n <= 0) return 0;
  const unsigned b = (unsigned)std::min<long long>((n + 255) / 256, 8192);
  if (es == 2)
    hipLaunchKernelGGL(k_box_move<int16_t>, dim3(b), dim3(256), 0, (hipStream_t)stream,
                       (int16_t *)grid, (int16_t *)packed, ext[0], ext[1], ext[2], px, pxy,
                       xoff, to_grid);
  else
    hipLaunchKernelGGL(k_box_move<float>, dim3(b), dim3(256), 0, (hipStream_t)stream,
                       (float *)grid, (float *)packed, ext[0], ext[1], ext[2], px, pxy, xoff,
                       to_grid);
  return (int)hipGetLastError();
}

// One block per tile of a batch of flat one-cell tiles: does every cell the
// tile owns (its run of columns of the x-y plane, its z chunk) hold the same
// medium in all six media arrays?  out[t]: that medium, else -1.
__global__ void __launch_bounds__(256) k_batch_tile_media(BatchDev bd, int *out) {
  const int t = blockIdx.x;
  int lo = 0, hi = bd.ngrid - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (bd.start[mid] <= t) lo = mid; else hi = mid - 1;
  }
  const DevGrid &g = bd.grids[lo];
  const TileMap &tm = bd.tms[lo];
  if (!tm.flat) {
    if (threadIdx.x == 0) out[t] = -1;
    return;
  }
  const int blk = t - bd.start[lo];
  const int bx = blk % tm.tx, bz = blk / (tm.tx * tm.ty);
  const long long cols = (long long)(g.nx + 1) * (g.ny + 1);
  const long long f0 = (long long)bx * 64 * tm.bdy;
  const long long f1 = f0 + 64LL * tm.bdy < cols ? f0 + 64LL * tm.bdy : cols;
  const int k0 = bz * tm.kc, k1 = min(k0 + tm.kc, g.nz + 1);
  const long long X = f1 - f0, Z = k1 - k0;
  __shared__ int bad;
  const int i0 = (int)(f0 % (g.nx + 1)), j0 = (int)(f0 / (g.nx + 1));
  const int m0 = g.mi[0][fidx(g, i0, j0, k0)];
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  int b = 0;
  for (long long q = threadIdx.x; q < X * Z && !b; q += blockDim.x) {
    const long long f = f0 + q % X;
    const int i = (int)(f % (g.nx + 1)), j = (int)(f / (g.nx + 1)), k = k0 + (int)(q / X);
    const long long p = fidx(g, i, j, k);
#pragma unroll
    for (int c = 0; c < 6; ++c) b |= g.mi[c][p] != m0;
  }
  if (b) atomicOr(&bad, 1);
  __syncthreads();
  if (threadIdx.x == 0) out[t] = (bad || X <= 0 || Z <= 0) ? -1 : m0;
}

int launch_batch_tile_media(const BatchDev &bd, int *out, void *stream) {
  if (bd.total <= 0) return 0;
  hipLaunchKernelGGL(k_batch_tile_media, dim3((unsigned)bd.total), dim3(256), 0,
                     (hipStream_t)stream, bd, out);
  return (int)hipGetLastError();
}

// One block per step tile of a pole class without sources: the stored
// previous field of every dispersive cell the tile owns (E and H, inside the
// component's sweep and pole box) set to the cell's field, which is what the
// step's pole update left there when it skipped the store.
__global__ void __launch_bounds__(256) k_prev_settle(DevGrid g, const StepTile *tiles) {
  const StepTile t = tiles[blockIdx.x];
  const int x0 = t.x & 0xfffff, wx = t.x >> 20, y0 = t.y & 0xfffff, wy = t.y >> 20;
  const int k0 = t.k0, k1 = (int)((unsigned)t.k1 & kTileK1Mask);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const long long X = min(x0 + wx, nx + 1) - x0, Y = min(y0 + wy, ny + 1) - y0,
                  Z = min(k1, nz + 1) - k0;
  if (X <= 0 || Y <= 0 || Z <= 0) return;
  const long long n = X * Y * Z;
  for (long long q = threadIdx.x; q < n; q += blockDim.x) {
    const int i = x0 + (int)(q % X), j = y0 + (int)((q / X) % Y), k = k0 + (int)(q / (X * Y));
    const long long p = fidx(g, i, j, k);
    for (int side = 0; side < 2; ++side) {
      const AdeSet &A = g.ade[side];
      if (!A.has) continue;
      for (int c = 0; c < 3; ++c) {
        const bool on = side == 0 ? (c == 0 ? i < nx : c == 1 ? j < ny : k < nz)
                                  : (c == 0 ? (j < ny && k < nz)
                                            : c == 1 ? (i < nx && k < nz) : (i < nx && j < ny));
        const int *b = A.abox[c];
        if (!on || i < b[0] || i >= b[0] + b[3] || j < b[1] || j >= b[1] + b[4] || k < b[2] ||
            k >= b[2] + b[5])
          continue;
        const int comp = 3 * side + c;
        if (A.dslot[g.mi[comp][p]] < 0) continue;
        const long long zo = (long long)(k - b[2]) * b[3] * b[4];
        const long long axy = (long long)(j - b[1]) * b[3] + (i - b[0]);
        A.prev[c][zo + axy] = g.f[comp][p];
      }
    }
  }
}

int launch_prev_settle(const DevGrid &g, const StepTile *tiles, int n, void *stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_prev_settle, dim3((unsigned)n), dim3(256), 0, (hipStream_t)stream, g,
                     tiles);
  return (int)hipGetLastError();
}

// One block per four-cell step tile: does every cell it reads (owned, the
// halo column x0 + wx and row y0 + wy, planes k0..k1, inside the grid) hold
// the same medium in all six media arrays?  out: that medium, else -1.
__global__ void __launch_bounds__(256) k_tile_media(DevGrid g, const StepTile *tiles, int *out) {
  const StepTile t = tiles[blockIdx.x];
  const int x0 = t.x & 0xfffff, wx = t.x >> 20, y0 = t.y & 0xfffff, wy = t.y >> 20;
  const int k0 = t.k0, k1 = (int)((unsigned)t.k1 & kTileK1Mask);
  const int xe = min(x0 + wx, g.nx), ye = min(y0 + wy, g.ny), ke = min(k1, g.nz);
  const long long X = xe - x0 + 1, Y = ye - y0 + 1, Z = ke - k0 + 1;
  __shared__ int bad;
  const int m0 = g.mi[0][fidx(g, x0, y0, k0)];
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  const long long n = X * Y * Z;
  int b = 0;
  for (long long q = threadIdx.x; q < n && !b; q += blockDim.x) {
    const int x = x0 + (int)(q % X), y = y0 + (int)((q / X) % Y), z = k0 + (int)(q / (X * Y));
    const long long p = fidx(g, x, y, z);
#pragma unroll
    for (int c = 0; c < 6; ++c) b |= g.mi[c][p] != m0;
  }
  if (b) atomicOr(&bad, 1);
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = bad ? -1 : m0;
}

int launch_tile_media(const DevGrid &g, const StepTile *tiles, int n, int *out, void *stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_tile_media, dim3((unsigned)n), dim3(256), 0, (hipStream_t)stream, g,
                     tiles, out);
  return (int)hipGetLastError();
}

// The same box read out of a field array as fp32: binary16 values times the
// array's scale (one rounding, as the host conversion), fp32 values as they
// are.
template <class T>
__global__ void __launch_bounds__(256) k_box_gather_f32(const T *grid, float *packed, float scale,
                                                       long long ex, long long ey, long long ez,
                                                       long long px, long long pxy, int xoff) {
  const long long n = ex * ey * ez;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const long long i = t % ex, j = (t / ex) % ey, k = t / (ex * ey);
    const long long p = k * pxy + j * px + (xoff - 1) + i;
    if constexpr (sizeof(T) == 2)
      packed[t] = __half2float(grid[p]) * scale;
    else
      packed[t] = grid[p];
  }
}

int launch_box_gather_f32(const void *grid, size_t es, float scale, float *packed,
                          const long long ext[3], long long px, long long pxy, int xoff,
                          void *stream) {
  const long long n = ext[0] * ext[1] * ext[2];
  if (n <= 0) return 0;
  const unsigned b = (unsigned)std::min<long long>((n + 255) / 256, 8192);
  if (es == 2)
    hipLaunchKernelGGL(k_box_gather_f32<__half>, dim3(b), dim3(256), 0, (hipStream_t)stream,
                       (const __half *)grid, packed, scale, ext[0], ext[1], ext[2], px, pxy,
                       xoff);
  else
    hipLaunchKernelGGL(k_box_gather_f32<float>, dim3(b), dim3(256), 0, (hipStream_t)stream,
                       (const float *)grid, packed, scale, ext[0], ext[1], ext[2], px, pxy,
                       xoff);
  return (int)hipGetLastError();
}

// 8-bit copy of a media index array (every index below 256)
__global__ void __launch_bounds__(256) k_media_narrow(const int16_t *src, uint8_t *dst,
                                                     long long n) {
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x)
    dst[t] = (uint8_t)src[t];
}

int launch_media_narrow(const int16_t *src, uint8_t *dst, long long n, void *stream) {
  if (n <= 0) return 0;
  const unsigned b = (unsigned)std::min<long long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_media_narrow, dim3(b), dim3(256), 0, (hipStream_t)stream, src, dst, n);
  return (int)hipGetLastError();
}

// Incid of every TF/SF face cell for the one-launch step (the same
// arithmetic as k_tfsf_list, which applies it in place)
__device__ __forceinline__ void pw_incident_face(const PwFace *faces, int nf0, const StepSrc *sp,
                                                 const int *dn, float *inc, int q) {
  const PwFace &f = faces[q];
  const int half = q >= nf0 ? 1 : 0;
  const int ua = f.axis == 0 ? 1 : 0;
  const int va = f.axis == 2 ? 1 : 2;
  const long long nu = f.hi[ua] - f.lo[ua] + 1, nv = f.hi[va] - f.lo[va] + 1;
  const StepSrc &S = *sp;
  const int n = *dn;
  const float tm = half ? (float)(S.tiempo[n] + (double)0.5f * S.dt) : (float)S.tiempo[n];
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < nu * nv;
       c += (long long)gridDim.x * blockDim.x) {
    int cc[3];
    cc[f.axis] = f.fixed;
    cc[ua] = f.lo[ua] + (int)(c % nu);
    cc[va] = f.lo[va] + (int)(c / nu);
    inc[f.inc + c] = pw_incid(f, S.waves, S.phys, S.pstride, tm, cc);
  }
}

__global__ void __launch_bounds__(256) k_pw_incident(const PwFace *faces, int nf0,
                                                     const StepSrc *sp, const int *dn,
                                                     float *inc) {
  pw_incident_face(faces, nf0, sp, dn, inc, blockIdx.y);
}

// the same for the grids of a batch: blockIdx.z the grid, blockIdx.y its face
__global__ void __launch_bounds__(256) k_pw_incident_many(const PwIncJob *jobs) {
  const PwIncJob &J = jobs[blockIdx.z];
  if ((int)blockIdx.y >= J.nf) return;
  pw_incident_face(J.faces, J.nf0, J.src, J.dn, J.inc, blockIdx.y);
}

int launch_pw_incident(const PwFace *faces, int nf0, int nf, long long maxcells,
                       const StepSrc *src, const int *dn, float *inc, void *stream) {
  if (nf <= 0 || maxcells <= 0) return 0;
  const long long b = std::min<long long>((maxcells + 255) / 256, 4096);
  hipLaunchKernelGGL(k_pw_incident, dim3((unsigned)b, (unsigned)nf), dim3(256), 0,
                     (hipStream_t)stream, faces, nf0, src, dn, inc);
  return (int)hipGetLastError();
}

int launch_pw_incident_many(const PwIncJob *jobs, int njobs, int maxnf, long long maxcells,
                            void *stream) {
  if (njobs <= 0 || maxnf <= 0 || maxcells <= 0) return 0;
  const long long b = std::min<long long>((maxcells + 255) / 256, 1024);
  hipLaunchKernelGGL(k_pw_incident_many, dim3((unsigned)b, (unsigned)maxnf, (unsigned)njobs),
                     dim3(256), 0, (hipStream_t)stream, jobs);
  return (int)hipGetLastError();
}

__global__ void k_tick_many(int *const *counters, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) *counters[t] += 1;
}

int launch_tick_many(int *const *counters, int n, void *stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_tick_many, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                     counters, n);
  return (int)hipGetLastError();
}

// Frequency-domain sums of one box at step n (or *dn), observation.F90:
// 3553-3600: one thread per (point, frequency); the three components of the
// point are read once per thread and multiplied by the complex(8) phase
// factor of the step (each part scaled by the real sample).
__global__ void __launch_bounds__(256) k_dft(DevGrid g, DftDev d, int n, const int *dn) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long npt = (long long)d.cnt[0] * d.cnt[1] * d.cnt[2];
  if (t >= npt * d.nfreq) return;
  if (dn) n = *dn;
  const long long pt = t / d.nfreq;
  const int f = (int)(t - pt * d.nfreq);
  const int it = (int)(pt % d.cnt[0]);
  const int jt = (int)((pt / d.cnt[0]) % d.cnt[1]);
  const int kt = (int)(pt / ((long long)d.cnt[0] * d.cnt[1]));
  const long long p = fidx(g, d.first[0] + it * d.stride[0], d.first[1] + jt * d.stride[1],
                           d.first[2] + kt * d.stride[2]);
  const double2 a = d.aux[(long long)n * d.nfreq + f];
  for (int c = 0; c < 3; ++c) {
    const double v = (double)field_ld(g, 3 * d.kind + c, p);
    double2 *acc = d.acc + (pt * 3 + c) * d.nfreq + f;
    const double2 o = *acc;
    *acc = make_double2(o.x + a.x * v, o.y + a.y * v);
  }
}

int launch_dft(const DevGrid &g, const DftDev &d, int n, const int *dn, void *stream) {
  const long long work = (long long)d.cnt[0] * d.cnt[1] * d.cnt[2] * d.nfreq;
  if (work <= 0) return 0;
  const unsigned blocks = (unsigned)((work + 255) / 256);
  hipLaunchKernelGGL(k_dft, dim3(blocks), dim3(256), 0, (hipStream_t)stream, g, d, n, dn);
  return (int)hipGetLastError();
}

// UpdateObservation point probes (observation.F90:2620-2672): one row of
// samples per call; *count advances on the device so captured graphs can
// replay it.
__global__ void k_probes(DevGrid g, const ProbeDev *pr, int np, float *buf,
                         int *count, int cap) {
  __shared__ int row;
  if (threadIdx.x == 0) row = *count;
  __syncthreads();
  if (row < cap) {
    for (int t = threadIdx.x; t < np; t += blockDim.x)
      buf[(long long)row * np + t] = field_ld(g, pr[t].comp, pr[t].p);
  }
  __syncthreads();
  if (threadIdx.x == 0 && row < cap) *count = row + 1;
}

// one block per probe set, as k_probes
__global__ void k_probes_many(const DevGrid *grids, const ProbeSet *sets) {
  const ProbeSet &ps = sets[blockIdx.x];
  const DevGrid &g = grids[ps.grid];
  __shared__ int row;
  if (threadIdx.x == 0) row = *ps.count;
  __syncthreads();
  if (row < ps.cap) {
    for (int t = threadIdx.x; t < ps.np; t += blockDim.x)
      ps.buf[(long long)row * ps.np + t] = field_ld(g, ps.pr[t].comp, ps.pr[t].p);
  }
  __syncthreads();
  if (threadIdx.x == 0 && row < ps.cap) *ps.count = row + 1;
}

int launch_probes_many(const DevGrid *grids, const ProbeSet *sets, int nsets, void *stream) {
  if (nsets <= 0) return 0;
  hipLaunchKernelGGL(k_probes_many, dim3(nsets), dim3(256), 0, (hipStream_t)stream, grids,
                     sets);
  return (int)hipGetLastError();
}

int launch_probes(const DevGrid &g, const ProbeDev *pr, int np, float *buf,
                  int *count, int cap, void *stream) {
  if (np <= 0) return 0;
  hipLaunchKernelGGL(k_probes, dim3(1), dim3(256), 0, (hipStream_t)stream, g, pr, np,
                     buf, count, cap);
  return (int)hipGetLastError();
}

// bounding box of the dispersive edges of component comp over its sweep
// (the node lists InitEDispersives and InitMdispersives gather,
// electricdispersive.F90:148-244, magneticdispersive.F90:139-265).
// Each thread keeps a running min/max, the block reduces in LDS and one lane
// folds the block result into out[] with global atomics.
__global__ void k_disp_bbox(DevGrid g, int comp, int xlast, int ylast, int zlast,
                            int32_t *out) {
  __shared__ int red[7][256];
  int lo[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff};
  int hi[3] = {-0x7fffffff - 1, -0x7fffffff - 1, -0x7fffffff - 1};
  int cnt = 0;
  const long long nxr = xlast + 1, nyr = ylast + 1;
  const long long tot = nxr * nyr * (zlast + 1);
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < tot;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t % nxr);
    const int j = (int)((t / nxr) % nyr);
    const int k = (int)(t / (nxr * nyr));
    if (g.ade[comp / 3].dslot[g.mi[comp][fidx(g, i, j, k)]] >= 0) {
      lo[0] = min(lo[0], i); lo[1] = min(lo[1], j); lo[2] = min(lo[2], k);
      hi[0] = max(hi[0], i); hi[1] = max(hi[1], j); hi[2] = max(hi[2], k);
      ++cnt;
    }
  }
  for (int a = 0; a < 3; ++a) {
    red[a][threadIdx.x] = lo[a];
    red[3 + a][threadIdx.x] = hi[a];
  }
  red[6][threadIdx.x] = cnt;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      for (int a = 0; a < 3; ++a) {
        red[a][threadIdx.x] = min(red[a][threadIdx.x], red[a][threadIdx.x + w]);
        red[3 + a][threadIdx.x] = max(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + w]);
      }
      red[6][threadIdx.x] += red[6][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    for (int a = 0; a < 3; ++a) {
      atomicMin(&out[a], red[a][0]);
      atomicMax(&out[3 + a], red[3 + a][0]);
    }
    atomicAdd(&out[6], red[6][0]);
  }
}

// side 0: E component c (sweep n-1 along c, n across); side 1: H component
// c (n along c, n-1 across)
int launch_disp_bbox(const DevGrid &g, int side, int c, int32_t *out, void *stream) {
  const int o = side ? 0 : 1, a = side ? 1 : 0;
  const int xl = c == 0 ? g.nx - o : g.nx - a;
  const int yl = c == 1 ? g.ny - o : g.ny - a;
  const int zl = c == 2 ? g.nz - o : g.nz - a;
  hipLaunchKernelGGL(k_disp_bbox, dim3(2048), dim3(256), 0, (hipStream_t)stream, g,
                     3 * side + c, xl, yl, zl, out);
  return (int)hipGetLastError();
}

// ---- first-order Mur, AdvanceMagneticMUR (bordersmur.F90:1129-1600) -----
// One thread per face cell: H_ghost = Past_inner + CAB1[m] (H_inner -
// Past_ghost), m the inner cell's media index, then both Past planes take the
// new field.  Ghost cells of different faces never coincide and no face reads
// another's ghost plane, so one launch equals the reference's face-by-face
// loops followed by its Past updates.  EXACT: no contraction (the file is
// built without it); FAST: one fused multiply-add.
template <bool EXACT>
__global__ void k_mur(DevGrid g, const MurRun *runs, int nrun, long long total) {
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    int r = 0;
    while (r + 1 < nrun && runs[r + 1].start <= t) ++r;
    const MurRun &m = runs[r];
    const long long q = t - m.start;
    const int u = (int)(q % m.nu), v = (int)(q / m.nu);
    int cb[3], ci[3];
    cb[m.ua] = ci[m.ua] = u;
    cb[m.va] = ci[m.va] = v;
    cb[m.axis] = m.ghost;
    ci[m.axis] = m.inner;
    const long long pb = fidx(g, cb[0], cb[1], cb[2]);
    const long long pi = fidx(g, ci[0], ci[1], ci[2]);
    const long long plane = (long long)m.nu * m.nv;
    const float past_b = m.past[q], past_i = m.past[plane + q];
    const float h_i = field_ld(g, m.comp, pi);
    const float c = m.cab1[g.mi[m.comp][pi]];
    const float h_b = EXACT ? past_i + c * (h_i - past_b) : __builtin_fmaf(c, h_i - past_b, past_i);
    field_st(g, m.comp, pb, h_b);
    m.past[q] = field_ld(g, m.comp, pb);
    m.past[plane + q] = h_i;
  }
}

int launch_mur(const DevGrid &g, const MurRun *runs, int nrun, long long total, int exact,
               void *stream) {
  if (total <= 0) return 0;
  const long long want = (total + 255) / 256;
  const dim3 grid((unsigned)(want < 4096 ? want : 4096));
  if (exact) hipLaunchKernelGGL(k_mur<true>, grid, dim3(256), 0, (hipStream_t)stream, g, runs, nrun, total);
  else hipLaunchKernelGGL(k_mur<false>, grid, dim3(256), 0, (hipStream_t)stream, g, runs, nrun, total);
  return (int)hipGetLastError();
}

__global__ void k_tick(int *dn) { *dn = *dn + 1; }

int launch_tick(int *dn, void *stream) {
  hipLaunchKernelGGL(k_tick, dim3(1), dim3(1), 0, (hipStream_t)stream, dn);
  return (int)hipGetLastError();
}


__global__ void k_fill_i16(int16_t *a, long long n, int16_t v) {
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x)
    a[t] = v;
}

int launch_fill_i16(int16_t *a, long long n, int16_t v, void *stream) {
  hipLaunchKernelGGL(k_fill_i16, dim3(4096), dim3(256), 0, (hipStream_t)stream, a, n, v);
  return (int)hipGetLastError();
}

// PEC outer walls: tangential E and normal H of a face get media 0
// (healer.F90:1442-1617).  One thread per (u, v) position of a face plane.
__global__ void k_pec_walls(DevGrid g, int16_t *m0, int16_t *m1, int16_t *m2,
                            int16_t *m3, int16_t *m4, int16_t *m5, int faces) {
  int16_t *mi[6] = {m0, m1, m2, m3, m4, m5};
  const int n[3] = {g.nx, g.ny, g.nz};
  const int face = blockIdx.y;  // 0..5
  if (!(faces & (1 << face))) return;
  const int a = face >> 1;
  const int v = (face & 1) ? n[a] : 0;
  const int b = (a + 1) % 3, c = (a + 2) % 3;
  const long long cnt = (long long)(n[b] + 1) * (n[c] + 1);
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < cnt;
       t += (long long)gridDim.x * blockDim.x) {
    int crd[3];
    crd[a] = v;
    crd[b] = (int)(t % (n[b] + 1));
    crd[c] = (int)(t / (n[b] + 1));
    const long long p = fidx(g, crd[0], crd[1], crd[2]);
    for (int comp = 0; comp < 3; ++comp)
      if (comp != a) mi[comp][p] = 0;
    mi[3 + a][p] = 0;
  }
}

int launch_pec_walls(const DevGrid &g, int16_t *mi[6], int faces, void *stream) {
  hipLaunchKernelGGL(k_pec_walls, dim3(1024, 6), dim3(256), 0, (hipStream_t)stream, g,
                     mi[0], mi[1], mi[2], mi[3], mi[4], mi[5], faces);
  return (int)hipGetLastError();
}

// CreateVolumeMM (healer.F90:66-234) for a voxelised shape: a voxel claims
// its 12 edges (priority strictly larger than the edge's medium) and its 6
// faces (larger or equal).  shape 0: sphere {cx,cy,cz,r} (voxel centre
// inside), 1: box {x0,x1,y0,y1,z0,z1}.  par[6..11] = launch box origin/size.
__device__ __forceinline__ bool voxel_in(const DevGrid &g, int shape, const float *par,
                                         int i, int j, int k) {
  if (i < 0 || j < 0 || k < 0 || i >= g.nx || j >= g.ny || k >= g.nz) return false;
  if (shape == 0) {
    const float dx = (float)i + 0.5f - par[0], dy = (float)j + 0.5f - par[1],
                dz = (float)k + 0.5f - par[2];
    return dx * dx + dy * dy + dz * dz <= par[3] * par[3];
  }
  return i >= (int)par[0] && i < (int)par[1] && j >= (int)par[2] && j < (int)par[3] &&
         k >= (int)par[4] && k < (int)par[5];
}

struct PaintArgs {
  int16_t *mi[6];
  float par[12];
  int shape;
  int16_t medium;
  int nprio;
  int ox, oy, oz, sx, sy, sz;
};

__global__ void k_paint(DevGrid g, PaintArgs a, const int32_t *prio) {
  const long long tot = (long long)a.sx * a.sy * a.sz;
  const int pn = prio[a.medium];
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < tot;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = a.ox + (int)(t % a.sx);
    const int j = a.oy + (int)((t / a.sx) % a.sy);
    const int k = a.oz + (int)(t / ((long long)a.sx * a.sy));
    if (i > g.nx || j > g.ny || k > g.nz || i < 0 || j < 0 || k < 0) continue;
    const long long p = fidx(g, i, j, k);
    // media outside the priority table are never overwritten
#define PR(c) (a.mi[c][p] < a.nprio ? prio[a.mi[c][p]] : 0x7fffffff)
#define IN(di, dj, dk) voxel_in(g, a.shape, a.par, i + (di), j + (dj), k + (dk))
    const bool v000 = IN(0, 0, 0);
    // Ex(i,j,k): voxels (i, j-1..j, k-1..k)
    if (i < g.nx && (v000 || IN(0, -1, 0) || IN(0, 0, -1) || IN(0, -1, -1)))
      if (pn > PR(0)) a.mi[0][p] = a.medium;
    // Ey: (i-1..i, j, k-1..k)
    if (j < g.ny && (v000 || IN(-1, 0, 0) || IN(0, 0, -1) || IN(-1, 0, -1)))
      if (pn > PR(1)) a.mi[1][p] = a.medium;
    // Ez: (i-1..i, j-1..j, k)
    if (k < g.nz && (v000 || IN(-1, 0, 0) || IN(0, -1, 0) || IN(-1, -1, 0)))
      if (pn > PR(2)) a.mi[2][p] = a.medium;
    // Hx: (i-1..i, j, k)
    if (j < g.ny && k < g.nz && (v000 || IN(-1, 0, 0)))
      if (pn >= PR(3)) a.mi[3][p] = a.medium;
    // Hy: (i, j-1..j, k)
    if (i < g.nx && k < g.nz && (v000 || IN(0, -1, 0)))
      if (pn >= PR(4)) a.mi[4][p] = a.medium;
    // Hz: (i, j, k-1..k)
    if (i < g.nx && j < g.ny && (v000 || IN(0, 0, -1)))
      if (pn >= PR(5)) a.mi[5][p] = a.medium;
#undef IN
#undef PR
  }
}

int launch_paint(const DevGrid &g, int16_t *mi[6], int shape, const float *par,
                 int16_t medium, const int32_t *prio, int nprio, void *stream) {
  PaintArgs a;
  a.nprio = nprio;
  for (int c = 0; c < 6; ++c) a.mi[c] = mi[c];
  for (int t = 0; t < 12; ++t) a.par[t] = par[t];
  a.shape = shape;
  a.medium = medium;
  a.ox = (int)par[6]; a.oy = (int)par[7]; a.oz = (int)par[8];
  a.sx = (int)par[9]; a.sy = (int)par[10]; a.sz = (int)par[11];
  if (a.sx <= 0 || a.sy <= 0 || a.sz <= 0) return 0;
  hipLaunchKernelGGL(k_paint, dim3(8192), dim3(256), 0, (hipStream_t)stream, g, a, prio);
  return (int)hipGetLastError();
}

}  // namespace sfd
