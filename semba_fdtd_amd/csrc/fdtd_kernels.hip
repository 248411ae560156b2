// fdtd_kernels.hip -- the two half-step passes of the Yee time loop on gfx950.
//
// This file is compiled twice: with -DSFD_NS=fast (FMA contraction allowed,
// conjugate pole pairs folded into one fp32 complex state) and with
// -DSFD_NS=exact -DSFD_EXACT=1 -ffp-contract=off (complex(8) pole state,
// one rounding per operation in the reference's expression order).
//
// E pass  = Advance_Ex/Ey/Ez (timestepping.F90:2165-2285)
//         + AdvanceelectricCPML (borderscpml.F90:584-876)
//         + AdvanceEDispersiveE (electricdispersive.F90:304-362)
// H pass  = Advance_Hx/Hy/Hz (timestepping.F90:2294-2409)
//         + AdvanceMagneticCPML (borderscpml.F90:881-1176)
// fused per edge, in the reference's per-edge order (Yee, then the psi
// corrections in region order y-slab, z-slab, x-slab, then the poles), so
// each field value is read and written once per half step.
//
// Work decomposition: a block of 64 x TY threads owns a 64-wide x segment of
// TY rows and marches a chunk of z planes.  Each thread carries the previous
// (E pass) or next (H pass) plane of the two z-differenced components in
// registers, so the z neighbour costs no extra HBM read; the x and y
// neighbours come from the same or an adjacent 128-byte line and hit L1/L2.
// Loads of a wave are 64 consecutive x: 256 B of fp32, 128 B of int16.
// The thread domain covers [0,nx]x[0,ny]x[0,nz]; each component updates the
// part of it that is inside its own sweep.
#include <hip/hip_runtime.h>
#include "fdtd_internal.h"

#ifndef SFD_NS
#error "compile with -DSFD_NS=fast or -DSFD_NS=exact"
#endif
#ifndef SFD_EXACT
#define SFD_EXACT 0
#endif

namespace sfd {
namespace SFD_NS {

__device__ __forceinline__ void decode_tile(const TileMap &tm, int &bx, int &by,
                                            int &bz) {
  int b = blockIdx.x;
  if (tm.remap) {
    // hardware dispatches block b to XCD b % 8; give each XCD one contiguous
    // range of tiles so that x/y neighbours share that XCD's L2.
    const int q = tm.nblk >> 3, r = tm.nblk & 7;
    const int xcd = b & 7, loc = b >> 3;
    b = (xcd < r) ? xcd * (q + 1) + loc : r * (q + 1) + (xcd - r) * q + loc;
  }
  bx = b % tm.tx;
  b /= tm.tx;
  by = b % tm.ty;
  bz = b / tm.ty;
}

// slab coordinate along a stretched axis.  "node" axes hold n+1 positions
// (tangential E, normal H), "cell" axes n.  High slabs start after the
// interior: node n-Lhi+1, cell n-Lhi (borderscpml.F90:149-193).
__device__ __forceinline__ bool slab_node(int c, int n, int L0, int L1, int &q) {
  if (c < L0) { q = c; return true; }
  const int hs = n - L1 + 1;
  if (c >= hs) { q = L0 + (c - hs); return true; }
  return false;
}
__device__ __forceinline__ bool slab_cell(int c, int n, int L0, int L1, int &q) {
  if (c < L0) { q = c; return true; }
  const int hs = n - L1;
  if (c >= hs) { q = L0 + (c - hs); return true; }
  return false;
}

// psi <- b*psi + d*c ; returns the new psi
__device__ __forceinline__ float psi_step(float *psi, long long s, float b,
                                          float d, float c) {
  const float v = b * psi[s] + d * c;
  psi[s] = v;
  return v;
}

// Pole-residue update of one dispersive E edge (electricdispersive.F90:
// 314-360): E -= sum_p Re(G3_p J_p); J_p = kappa_p J_p + beta_p (E-Eprev)/dt;
// Eprev = E.
template <int C>
__device__ __forceinline__ float ade_edge(const DevGrid &g, float e, int slot,
                                          long long s) {
  const int np = g.anp[slot];
  const long long N = g.nstate[C];
  float *ep = g.eprev[C];
#if SFD_EXACT
  const double *cf = g.aded + (size_t)slot * kPMax * 6;
  double2 *J = reinterpret_cast<double2 *>(g.jst[C]);
  double2 jv[kPMax];
#pragma unroll
  for (int q = 0; q < kPMax; ++q) {
    if (q < np) {
      jv[q] = J[q * N + s];
      const double re = cf[q * 6 + 4] * jv[q].x - cf[q * 6 + 5] * jv[q].y;
      e = (float)((double)e - re);
    }
  }
  const float diff = e - ep[s];
#pragma unroll
  for (int q = 0; q < kPMax; ++q) {
    if (q < np) {
      const double kr = cf[q * 6 + 0], ki = cf[q * 6 + 1];
      const double ar = kr * jv[q].x - ki * jv[q].y;
      const double ai = kr * jv[q].y + ki * jv[q].x;
      const double br = cf[q * 6 + 2] * (double)diff / g.dt;
      const double bi = cf[q * 6 + 3] * (double)diff / g.dt;
      J[q * N + s] = make_double2(ar + br, ai + bi);
    }
  }
#else
  const float *cf = g.adef + (size_t)slot * kPMax * 6;
  float2 *J = reinterpret_cast<float2 *>(g.jst[C]);
  float2 jv[kPMax];
  float sum = 0.0f;
#pragma unroll
  for (int q = 0; q < kPMax; ++q) {
    if (q < np) {
      jv[q] = J[q * N + s];
      sum += cf[q * 6 + 4] * jv[q].x - cf[q * 6 + 5] * jv[q].y;
    }
  }
  e = e - sum;
  const float diff = e - ep[s];
#pragma unroll
  for (int q = 0; q < kPMax; ++q) {
    if (q < np) {
      const float kr = cf[q * 6 + 0], ki = cf[q * 6 + 1];
      J[q * N + s] = make_float2(kr * jv[q].x - ki * jv[q].y + cf[q * 6 + 2] * diff,
                                 kr * jv[q].y + ki * jv[q].x + cf[q * 6 + 3] * diff);
    }
  }
#endif
  ep[s] = e;
  return e;
}

// wave-cooperative ADE for component C: lanes whose edge is dispersive find
// their compact slot as segment base + rank among the set ballot bits.
template <int C>
__device__ __forceinline__ float ade_wave(const DevGrid &g, bool valid, int m,
                                         float e, long long row, int seg,
                                         int lane, bool &touched) {
  const int slot = valid ? g.dslot[m] : -1;
  const unsigned long long mask = __ballot(slot >= 0);
  if (mask) {
    if (slot >= 0) {
      const long long base = g.segoff[C][row * g.nseg + seg];
      const int rank = __popcll(mask & ((1ull << lane) - 1ull));
      e = ade_edge<C>(g, e, slot, base + rank);
      touched = true;
    }
  }
  return e;
}

template <bool YEE, bool CPML, bool ADE>
__global__ void __launch_bounds__(256) k_epass(DevGrid g, TileMap tm) {
  int bx, by, bz;
  decode_tile(tm, bx, by, bz);
  const int lane = threadIdx.x;
  const int i = bx * kWave + lane;
  const int j = by * blockDim.y + threadIdx.y;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const bool act = (i <= nx) && (j <= ny);
  const int k0 = bz * tm.kc;
  const int k1 = min(k0 + tm.kc, nz + 1);
  const long long px = g.px, pxy = g.pxy;
  long long p = (long long)(k0 + 1) * pxy + (long long)(j + 1) * px + kXOff + i;
  float *__restrict__ Ex = g.f[0];
  float *__restrict__ Ey = g.f[1];
  float *__restrict__ Ez = g.f[2];
  const float *__restrict__ Hx = g.f[3];
  const float *__restrict__ Hy = g.f[4];
  const float *__restrict__ Hz = g.f[5];
  constexpr bool NEEDH = YEE || CPML;
  float hx_km = 0.0f, hy_km = 0.0f, idyh = 0.0f, idxh = 0.0f;
  if (act) {
    if (NEEDH) {
      hx_km = Hx[p - pxy];
      hy_km = Hy[p - pxy];
    }
    idyh = g.id[4][j + 1];
    idxh = g.id[3][i + 1];
  }
  // CPML slab membership along x and y is fixed for the thread
  int qxn = 0, qyn = 0;
  bool sxn = false, syn = false;
  if (CPML && act) {
    sxn = slab_node(i, nx, g.L[0][0], g.L[0][1], qxn);
    syn = slab_node(j, ny, g.L[1][0], g.L[1][1], qyn);
  }
  for (int k = k0; k < k1; ++k, p += pxy) {
    const bool vx = act && i < nx, vy = act && j < ny, vz = act && k < nz;
    float exn = 0.0f, eyn = 0.0f, ezn = 0.0f;
    int mx = 0, my = 0, mz = 0;
    float hx0 = 0.0f, hy0 = 0.0f;
    if (act) {
      const float idzh = g.id[5][k + 1];
      float hz0 = 0.0f, hz_jm = 0.0f, hx_jm = 0.0f, hz_im = 0.0f, hy_im = 0.0f;
      if (NEEDH) {
        hx0 = Hx[p]; hy0 = Hy[p]; hz0 = Hz[p];
        hz_jm = Hz[p - px]; hx_jm = Hx[p - px];
        hz_im = Hz[p - 1]; hy_im = Hy[p - 1];
      }
      int qzn = 0;
      const bool szn = CPML && slab_node(k, nz, g.L[2][0], g.L[2][1], qzn);
      if (vx) {
        mx = g.mi[0][p];
        float e = Ex[p];
        if (YEE)
          e = g.g1[mx] * e + g.g2[mx] * ((hz0 - hz_jm) * idyh - (hy0 - hy_km) * idzh);
        if (CPML) {
          const float g2 = g.g2[mx];
          if (syn) {  // psi_Exy, left/right (borderscpml.F90:621-623)
            const long long s = ((long long)k * (g.L[1][0] + g.L[1][1]) + qyn) * nx + i;
            e = e + g2 * psi_step(g.psi[PSI_EXY], s, g.pbe[1][j + 1], hz0 - hz_jm, g.pce[1][j + 1]);
          }
          if (szn) {  // psi_Exz, down/up (:712)
            const long long s = ((long long)qzn * (ny + 1) + j) * nx + i;
            e = e - g2 * psi_step(g.psi[PSI_EXZ], s, g.pbe[2][k + 1], hy0 - hy_km, g.pce[2][k + 1]);
          }
        }
        exn = e;
      }
      if (vy) {
        my = g.mi[1][p];
        float e = Ey[p];
        if (YEE)
          e = g.g1[my] * e + g.g2[my] * ((hx0 - hx_km) * idzh - (hz0 - hz_im) * idxh);
        if (CPML) {
          const float g2 = g.g2[my];
          if (szn) {  // psi_Eyz, down/up (:642... region down)
            const long long s = ((long long)qzn * ny + j) * (nx + 1) + i;
            e = e + g2 * psi_step(g.psi[PSI_EYZ], s, g.pbe[2][k + 1], hx0 - hx_km, g.pce[2][k + 1]);
          }
          if (sxn) {  // psi_Eyx, back/front
            const long long s = ((long long)k * ny + j) * (g.L[0][0] + g.L[0][1]) + qxn;
            e = e - g2 * psi_step(g.psi[PSI_EYX], s, g.pbe[0][i + 1], hz0 - hz_im, g.pce[0][i + 1]);
          }
        }
        eyn = e;
      }
      if (vz) {
        mz = g.mi[2][p];
        float e = Ez[p];
        if (YEE)
          e = g.g1[mz] * e + g.g2[mz] * ((hy0 - hy_im) * idxh - (hx0 - hx_jm) * idyh);
        if (CPML) {
          const float g2 = g.g2[mz];
          if (syn) {  // psi_Ezy, left/right
            const long long s = ((long long)k * (g.L[1][0] + g.L[1][1]) + qyn) * (nx + 1) + i;
            e = e - g2 * psi_step(g.psi[PSI_EZY], s, g.pbe[1][j + 1], hx0 - hx_jm, g.pce[1][j + 1]);
          }
          if (sxn) {  // psi_Ezx, back/front
            const long long s = ((long long)k * (ny + 1) + j) * (g.L[0][0] + g.L[0][1]) + qxn;
            e = e + g2 * psi_step(g.psi[PSI_EZX], s, g.pbe[0][i + 1], hy0 - hy_im, g.pce[0][i + 1]);
          }
        }
        ezn = e;
      }
    }
    bool tx = false, ty = false, tz = false;
    if (ADE) {
      if (!YEE && !CPML) {
        if (vx) { mx = g.mi[0][p]; exn = Ex[p]; }
        if (vy) { my = g.mi[1][p]; eyn = Ey[p]; }
        if (vz) { mz = g.mi[2][p]; ezn = Ez[p]; }
      }
      exn = ade_wave<0>(g, vx, mx, exn, (long long)k * (ny + 1) + j, bx, lane, tx);
      eyn = ade_wave<1>(g, vy, my, eyn, (long long)k * ny + j, bx, lane, ty);
      ezn = ade_wave<2>(g, vz, mz, ezn, (long long)k * (ny + 1) + j, bx, lane, tz);
    }
    constexpr bool ALL = YEE || CPML;
    if (vx && (ALL || tx)) Ex[p] = exn;
    if (vy && (ALL || ty)) Ey[p] = eyn;
    if (vz && (ALL || tz)) Ez[p] = ezn;
    if (NEEDH) { hx_km = hx0; hy_km = hy0; }
  }
}

template <bool YEE, bool CPML>
__global__ void __launch_bounds__(256) k_hpass(DevGrid g, TileMap tm) {
  int bx, by, bz;
  decode_tile(tm, bx, by, bz);
  const int lane = threadIdx.x;
  const int i = bx * kWave + lane;
  const int j = by * blockDim.y + threadIdx.y;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const bool act = (i <= nx) && (j <= ny);
  if (!act) return;  // no cross-lane operations in this pass
  const int k0 = bz * tm.kc;
  const int k1 = min(k0 + tm.kc, nz + 1);
  const long long px = g.px, pxy = g.pxy;
  long long p = (long long)(k0 + 1) * pxy + (long long)(j + 1) * px + kXOff + i;
  const float *__restrict__ Ex = g.f[0];
  const float *__restrict__ Ey = g.f[1];
  const float *__restrict__ Ez = g.f[2];
  float *__restrict__ Hx = g.f[3];
  float *__restrict__ Hy = g.f[4];
  float *__restrict__ Hz = g.f[5];
  const float idye = g.id[1][j + 1];
  const float idxe = g.id[0][i + 1];
  float ex_c = Ex[p], ey_c = Ey[p];
  int qxc = 0, qyc = 0;
  bool sxc = false, syc = false;
  if (CPML) {
    sxc = slab_cell(i, nx, g.L[0][0], g.L[0][1], qxc);
    syc = slab_cell(j, ny, g.L[1][0], g.L[1][1], qyc);
  }
  for (int k = k0; k < k1; ++k, p += pxy) {
    const float idze = g.id[2][k + 1];
    const float ex_kp = Ex[p + pxy], ey_kp = Ey[p + pxy];
    const float ez_c = Ez[p];
    const float ez_jp = Ez[p + px], ex_jp = Ex[p + px];
    const float ez_ip = Ez[p + 1], ey_ip = Ey[p + 1];
    int qzc = 0;
    const bool szc = CPML && slab_cell(k, nz, g.L[2][0], g.L[2][1], qzc);
    if (j < ny && k < nz) {  // Hx sweep: x[0,nx] y[0,ny-1] z[0,nz-1]
      const int m = g.mi[3][p];
      float h = Hx[p];
      if (YEE)
        h = g.gm1[m] * h + g.gm2[m] * ((ey_kp - ey_c) * idze - (ez_jp - ez_c) * idye);
      if (CPML) {
        const float gm2 = g.gm2[m];
        if (syc) {  // psi_Hxy (borderscpml.F90:921)
          const long long s = ((long long)k * (g.L[1][0] + g.L[1][1]) + qyc) * (nx + 1) + i;
          h = h - gm2 * psi_step(g.psi[PSI_HXY], s, g.pbm[1][j + 1], ez_jp - ez_c, g.pcm[1][j + 1]);
        }
        if (szc) {  // psi_Hxz (:1031)
          const long long s = ((long long)qzc * ny + j) * (nx + 1) + i;
          h = h + gm2 * psi_step(g.psi[PSI_HXZ], s, g.pbm[2][k + 1], ey_kp - ey_c, g.pcm[2][k + 1]);
        }
      }
      Hx[p] = h;
    }
    if (i < nx && k < nz) {  // Hy sweep: x[0,nx-1] y[0,ny] z[0,nz-1]
      const int m = g.mi[4][p];
      float h = Hy[p];
      if (YEE)
        h = g.gm1[m] * h + g.gm2[m] * ((ez_ip - ez_c) * idxe - (ex_kp - ex_c) * idze);
      if (CPML) {
        const float gm2 = g.gm2[m];
        if (szc) {  // psi_Hyz (:1011)
          const long long s = ((long long)qzc * (ny + 1) + j) * nx + i;
          h = h - gm2 * psi_step(g.psi[PSI_HYZ], s, g.pbm[2][k + 1], ex_kp - ex_c, g.pcm[2][k + 1]);
        }
        if (sxc) {  // psi_Hyx (:1121)
          const long long s = ((long long)k * (ny + 1) + j) * (g.L[0][0] + g.L[0][1]) + qxc;
          h = h + gm2 * psi_step(g.psi[PSI_HYX], s, g.pbm[0][i + 1], ez_ip - ez_c, g.pcm[0][i + 1]);
        }
      }
      Hy[p] = h;
    }
    if (i < nx && j < ny) {  // Hz sweep: x[0,nx-1] y[0,ny-1] z[0,nz]
      const int m = g.mi[5][p];
      float h = Hz[p];
      if (YEE)
        h = g.gm1[m] * h + g.gm2[m] * ((ex_jp - ex_c) * idye - (ey_ip - ey_c) * idxe);
      if (CPML) {
        const float gm2 = g.gm2[m];
        if (syc) {  // psi_Hzy (:941)
          const long long s = ((long long)k * (g.L[1][0] + g.L[1][1]) + qyc) * nx + i;
          h = h + gm2 * psi_step(g.psi[PSI_HZY], s, g.pbm[1][j + 1], ex_jp - ex_c, g.pcm[1][j + 1]);
        }
        if (sxc) {  // psi_Hzx (:1101)
          const long long s = ((long long)k * ny + j) * (g.L[0][0] + g.L[0][1]) + qxc;
          h = h - gm2 * psi_step(g.psi[PSI_HZX], s, g.pbm[0][i + 1], ey_ip - ey_c, g.pcm[0][i + 1]);
        }
      }
      Hz[p] = h;
    }
    ex_c = ex_kp;
    ey_c = ey_kp;
  }
}

static inline int grid_of(const TileMap &tm) { return tm.nblk; }

int launch_e(const DevGrid &g, const TileMap &tm, int yee, int cpml, int ade,
             void *stream) {
  const dim3 blk(kWave, (unsigned)(256 / kWave));
  hipStream_t s = (hipStream_t)stream;
  const int nb = grid_of(tm);
#define SFD_E(Y, C, A) hipLaunchKernelGGL((k_epass<Y, C, A>), dim3(nb), blk, 0, s, g, tm)
  if (yee && cpml && ade) SFD_E(true, true, true);
  else if (yee && cpml) SFD_E(true, true, false);
  else if (yee && ade) SFD_E(true, false, true);
  else if (yee) SFD_E(true, false, false);
  else if (cpml && ade) SFD_E(false, true, true);
  else if (cpml) SFD_E(false, true, false);
  else if (ade) SFD_E(false, false, true);
  else return 0;
#undef SFD_E
  return (int)hipGetLastError();
}

int launch_h(const DevGrid &g, const TileMap &tm, int yee, int cpml,
             void *stream) {
  const dim3 blk(kWave, (unsigned)(256 / kWave));
  hipStream_t s = (hipStream_t)stream;
  const int nb = grid_of(tm);
  if (yee && cpml) hipLaunchKernelGGL((k_hpass<true, true>), dim3(nb), blk, 0, s, g, tm);
  else if (yee) hipLaunchKernelGGL((k_hpass<true, false>), dim3(nb), blk, 0, s, g, tm);
  else if (cpml) hipLaunchKernelGGL((k_hpass<false, true>), dim3(nb), blk, 0, s, g, tm);
  else return 0;
  return (int)hipGetLastError();
}

}  // namespace SFD_NS
}  // namespace sfd
