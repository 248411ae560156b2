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
// The per-medium coefficients, the dispersive slot of each medium and the
// pole coefficients are staged in LDS once per block, so the media-index
// gather costs an LDS read instead of a dependent global load.  Each z step
// first issues every independent load (fields, media indices, then the pole
// state of dispersive edges) and only then computes, so the pole state's
// HBM latency overlaps the curl arithmetic.  The thread domain covers
// [0,nx]x[0,ny]x[0,nz]; each component updates the part inside its sweep.
#include <hip/hip_fp16.h>
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

// pole states prefetched into registers per edge: one Lorentz pair is two
// complex(8) states in EXACT and one folded complex64 state in FAST
#if SFD_EXACT
typedef double2 jstate_t;  // complex(8) pole state, reference layout
typedef double coef_t;
constexpr int kPre = 2;
#else
typedef float2 jstate_t;   // complex64 state, conjugate pairs folded
typedef float coef_t;
#ifndef SFD_PREFETCH_POLES
#define SFD_PREFETCH_POLES 1
#endif
constexpr int kPre = SFD_PREFETCH_POLES;
#endif

__device__ __forceinline__ void decode_tile(const TileMap &tm, int b, int &bx, int &by,
                                            int &bz) {
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

// The update arithmetic.  EXACT: plain products and sums (the file is built
// without contraction), the reference's order.  FAST: explicit fused
// multiply-adds, so that every variant of a pass (one or four cells per lane,
// one grid or a batch) rounds the same way and gives the same bits.
// a*ia - b*ib
__device__ __forceinline__ float curl2(float a, float ia, float b, float ib) {
#if SFD_EXACT
  return a * ia - b * ib;
#else
  return __builtin_fmaf(a, ia, -(b * ib));
#endif
}
// g1*f + g2*cu
__device__ __forceinline__ float yee_upd(float g1, float f, float g2, float cu) {
#if SFD_EXACT
  return g1 * f + g2 * cu;
#else
  return __builtin_fmaf(g1, f, g2 * cu);
#endif
}
// x + g*v
__device__ __forceinline__ float madd(float x, float g, float v) {
#if SFD_EXACT
  return x + g * v;
#else
  return __builtin_fmaf(g, v, x);
#endif
}

// psi <- b*psi + d*c ; returns the new psi
__device__ __forceinline__ float psi_step(float *psi, long long s, float b,
                                          float d, float c) {
  const float v = madd(d * c, b, psi[s]);
  psi[s] = v;
  return v;
}

// per-medium tables staged in LDS (TBL) or read from global memory
struct MedTab {
  const float2 *gg;     // {G1,G2} or {GM1,GM2}
  const int *slot;      // dispersive slot per medium (-1: not dispersive)
  const int *anp;       // stored poles per slot
  const coef_t *acf;    // per slot per pole {kr,ki,br,bi,g3r,g3i}
};

__device__ __forceinline__ const coef_t *ade_coefs(const AdeSet &ad) {
#if SFD_EXACT
  return ad.aded;
#else
  return ad.adef;
#endif
}

// LDS layout: gg[nm+1] | slot[nm+1] | anp[ndisp] | (8-byte aligned) acf.
// ad: the pole state whose tables are staged (ADE only).
template <bool TBL, bool ADE>
__device__ __forceinline__ MedTab stage_tables(const DevGrid &g, const float *a,
                                               const float *b, float2 *lds,
                                               const AdeSet *ad = nullptr) {
  MedTab t;
  const int nm = g.nmedia;
  if (TBL) {
    int *ls = reinterpret_cast<int *>(lds + (nm + 1));
    int *lnp = ls + (nm + 1);
    const int nd = ADE ? ad->ndisp : 0;
    const size_t off = ((size_t)(nm + 1) * 12 + (size_t)nd * 4 + 7) & ~(size_t)7;
    coef_t *lcf = reinterpret_cast<coef_t *>(reinterpret_cast<char *>(lds) + off);
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    const int nt = blockDim.x * blockDim.y;
    for (int m = tid; m <= nm; m += nt) {
      lds[m] = make_float2(a[m], b[m]);
      if (ADE) ls[m] = nd ? ad->dslot[m] : -1;
    }
    if (ADE) {
      const coef_t *cf = ade_coefs(*ad);
      for (int d = tid; d < nd; d += nt) lnp[d] = ad->anp[d];
      for (int q = tid; q < nd * kPMax * 6; q += nt) lcf[q] = cf[q];
    }
    __syncthreads();
    t.gg = lds;
    t.slot = ls;
    t.anp = lnp;
    t.acf = lcf;
  } else {
    t.gg = nullptr;
    t.slot = ADE ? ad->dslot : nullptr;
    t.anp = ADE ? ad->anp : nullptr;
    t.acf = ADE ? ade_coefs(*ad) : nullptr;
  }
  return t;
}

template <bool TBL>
__device__ __forceinline__ float2 coef(const MedTab &t, const float *a, const float *b, int m) {
  if (TBL) return t.gg[m];
  return make_float2(a[m], b[m]);
}

// ---- pole-residue ADE (electricdispersive.F90:304-362 on E,
// magneticdispersive.F90:306-362 on H, the same recursion) -----------------
// The pole state of every dispersive edge of component C lives in a dense
// x-fastest box (abox); s = the edge's slot in it.  AdePre holds the loaded
// state of one edge.
struct AdePre {
  long long s;
  int slot;
  int np;
  float ep;
  jstate_t j[kPre];
};

// F -= sum_p Re(G3_p J_p); J_p = kappa_p J_p + beta_p (F - Fprev)/dt;
// Fprev = F  (F = E with G3, or H with GM3).  The state is read from ai and
// the new state written to ao when wr (the same set for in-place updates).
// FAST arithmetic is written out as fused multiply-adds so that every kernel
// that inlines it rounds the same way.
template <int C>
__device__ __forceinline__ float ade_apply(const AdeSet &ai, const AdeSet &ao, bool wr,
                                           double dt, const MedTab &t, const AdePre &a,
                                           float e) {
  const coef_t *cf = t.acf + (size_t)a.slot * kPMax * 6;
  const long long N = ai.nstate[C];
  const jstate_t *Ji = reinterpret_cast<const jstate_t *>(ai.jst[C]);
  jstate_t *Jo = reinterpret_cast<jstate_t *>(ao.jst[C]);
#if SFD_EXACT
  // one pole at a time, rounding E after each (electricdispersive.F90:319-321)
#pragma unroll
  for (int q = 0; q < kPMax; ++q) {
    if (q >= a.np) break;
    const double2 jv = q < kPre ? a.j[q] : Ji[q * N + a.s];
    const double re = cf[q * 6 + 4] * jv.x - cf[q * 6 + 5] * jv.y;
    e = (float)((double)e - re);
  }
  const float diff = e - a.ep;
#pragma unroll
  for (int q = 0; q < kPMax; ++q) {
    if (q >= a.np) break;
    const double2 jv = q < kPre ? a.j[q] : Ji[q * N + a.s];
    const double kr = cf[q * 6 + 0], ki = cf[q * 6 + 1];
    const double kjr = kr * jv.x - ki * jv.y;
    const double kji = kr * jv.y + ki * jv.x;
    const double br = cf[q * 6 + 2] * (double)diff / dt;
    const double bi = cf[q * 6 + 3] * (double)diff / dt;
    if (wr) Jo[q * N + a.s] = make_double2(kjr + br, kji + bi);
  }
#else
  // folded fp32 form: beta is pre-divided by dt, G3 of a folded pair doubled
  (void)dt;
  float sum = 0.0f;
#pragma unroll
  for (int q = 0; q < kPMax; ++q) {
    if (q >= a.np) break;
    const float2 jv = q < kPre ? a.j[q] : Ji[q * N + a.s];
    sum = sum + __builtin_fmaf(cf[q * 6 + 4], jv.x, -(cf[q * 6 + 5] * jv.y));
  }
  e = e - sum;
  const float diff = e - a.ep;
#pragma unroll
  for (int q = 0; q < kPMax; ++q) {
    if (q >= a.np) break;
    const float2 jv = q < kPre ? a.j[q] : Ji[q * N + a.s];
    const float kr = cf[q * 6 + 0], ki = cf[q * 6 + 1];
    const float nr = __builtin_fmaf(kr, jv.x, __builtin_fmaf(-ki, jv.y, cf[q * 6 + 2] * diff));
    const float ni = __builtin_fmaf(kr, jv.y, __builtin_fmaf(ki, jv.x, cf[q * 6 + 3] * diff));
    if (wr) Jo[q * N + a.s] = make_float2(nr, ni);
  }
#endif
  if (wr) ao.prev[C][a.s] = e;
  return e;
}

// The dispersive-edge pass: AdvanceEDispersiveE (SIDE 0) or
// AdvanceMdispersiveH (SIDE 1) over the pole-state box of each component
// (blockIdx.y), after the Yee+CPML pass of that field.  Edges of the box are
// visited x-fastest, so the media index, field, previous-field and state
// loads of a wave are each one contiguous run; non-dispersive edges cost only
// their int16 media index.
template <bool TBL, int SIDE>
__global__ void __launch_bounds__(256) k_ade_box(DevGrid g) {
  extern __shared__ float2 lds_tab[];
  const AdeSet &ad = g.ade[SIDE];
  const MedTab tab = SIDE ? stage_tables<TBL, true>(g, g.gm1, g.gm2, lds_tab, &ad)
                          : stage_tables<TBL, true>(g, g.g1, g.g2, lds_tab, &ad);
  const int c = blockIdx.y;
  const int f = 3 * SIDE + c;   // field component
  const int *b = ad.abox[c];
  const long long bx = b[3], bxy = (long long)b[3] * b[4];
  const long long n = bxy * b[5];
  for (long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x; s < n;
       s += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(s % bx), y = (int)((s / bx) % b[4]), z = (int)(s / bxy);
    const long long p = (long long)(b[2] + z + 1) * g.pxy + (long long)(b[1] + y + 1) * g.px +
                        kXOff + b[0] + x;
    const int m = g.mi[f][p];
    const int slot = tab.slot[m];
    if (slot < 0) continue;
    AdePre a;
    a.s = s;
    a.slot = slot;
    a.np = tab.anp[slot];
    a.ep = ad.prev[c][s];
    const jstate_t *J = reinterpret_cast<const jstate_t *>(ad.jst[c]);
#pragma unroll
    for (int q = 0; q < kPre; ++q)
      if (q < a.np) a.j[q] = J[q * ad.nstate[c] + s];
    const float e = field_ld(g, f, p);
    float r;
    if (c == 0) r = ade_apply<0>(ad, ad, true, g.dt, tab, a, e);
    else if (c == 1) r = ade_apply<1>(ad, ad, true, g.dt, tab, a, e);
    else r = ade_apply<2>(ad, ad, true, g.dt, tab, a, e);
    field_st(g, f, p, r);
  }
}

// one tile (64 x blockDim.y lanes, one z chunk) of the E face pass
template <bool YEE, bool CPML, bool TBL>
__device__ __forceinline__ void epass_tile(const DevGrid &g, const TileMap &tm,
                                           const MedTab &tab, int blk) {
  int bx, by, bz;
  decode_tile(tm, blk, bx, by, bz);
  const int i = bx * kWave + threadIdx.x;
  const int j = by * blockDim.y + threadIdx.y;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (i > nx || j > ny) return;
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
  float hx_km = 0.0f, hy_km = 0.0f;
  if (NEEDH) {
    hx_km = Hx[p - pxy];
    hy_km = Hy[p - pxy];
  }
  const float idyh = g.id[4][j + 1];
  const float idxh = g.id[3][i + 1];
  const bool cx = i < nx, cy = j < ny;   // Ex needs i<nx, Ey j<ny, Ez k<nz
  // CPML slab membership along x and y is fixed for the thread
  int qxn = 0, qyn = 0;
  bool sxn = false, syn = false;
  if (CPML) {
    sxn = slab_node(i, nx, g.L[0][0], g.L[0][1], qxn);
    syn = slab_node(j, ny, g.L[1][0], g.L[1][1], qyn);
  }
  for (int k = k0; k < k1; ++k, p += pxy) {
    const bool cz = k < nz;
    // ---- issue the independent loads of this plane
    const float idzh = g.id[5][k + 1];
    float hx0 = 0.0f, hy0 = 0.0f, hz0 = 0.0f, hz_jm = 0.0f, hx_jm = 0.0f, hz_im = 0.0f,
          hy_im = 0.0f;
    if (NEEDH) {
      hx0 = Hx[p]; hy0 = Hy[p]; hz0 = Hz[p];
      hz_jm = Hz[p - px]; hx_jm = Hx[p - px];
      hz_im = Hz[p - 1]; hy_im = Hy[p - 1];
    }
    int mx = 0, my = 0, mz = 0;
    float ex = 0.0f, ey = 0.0f, ez = 0.0f;
    if (cx) { mx = g.mi[0][p]; ex = Ex[p]; }
    if (cy) { my = g.mi[1][p]; ey = Ey[p]; }
    if (cz) { mz = g.mi[2][p]; ez = Ez[p]; }
    int qzn = 0;
    const bool szn = CPML && slab_node(k, nz, g.L[2][0], g.L[2][1], qzn);
    // ---- compute
    if (cx) {
      float e = ex;
      const float2 c = coef<TBL>(tab, g.g1, g.g2, mx);
      if (YEE) e = yee_upd(c.x, e, c.y, curl2(hz0 - hz_jm, idyh, hy0 - hy_km, idzh));
      if (CPML) {
        if (syn) {  // psi_Exy, left/right (borderscpml.F90:609-623)
          const long long s = ((long long)k * (g.L[1][0] + g.L[1][1]) + qyn) * nx + i;
          e = madd(e, c.y, psi_step(g.psi[PSI_EXY], s, g.pbe[1][j + 1], hz0 - hz_jm, g.pce[1][j + 1]));
        }
        if (szn) {  // psi_Exz, down/up (:699-712)
          const long long s = ((long long)qzn * (ny + 1) + j) * nx + i;
          e = madd(e, -c.y, psi_step(g.psi[PSI_EXZ], s, g.pbe[2][k + 1], hy0 - hy_km, g.pce[2][k + 1]));
        }
      }
      Ex[p] = e;
    }
    if (cy) {
      float e = ey;
      const float2 c = coef<TBL>(tab, g.g1, g.g2, my);
      if (YEE) e = yee_upd(c.x, e, c.y, curl2(hx0 - hx_km, idzh, hz0 - hz_im, idxh));
      if (CPML) {
        if (szn) {  // psi_Eyz, down/up (:680-693)
          const long long s = ((long long)qzn * ny + j) * (nx + 1) + i;
          e = madd(e, c.y, psi_step(g.psi[PSI_EYZ], s, g.pbe[2][k + 1], hx0 - hx_km, g.pce[2][k + 1]));
        }
        if (sxn) {  // psi_Eyx, back/front (:790-802)
          const long long s = ((long long)k * ny + j) * (g.L[0][0] + g.L[0][1]) + qxn;
          e = madd(e, -c.y, psi_step(g.psi[PSI_EYX], s, g.pbe[0][i + 1], hz0 - hz_im, g.pce[0][i + 1]));
        }
      }
      Ey[p] = e;
    }
    if (cz) {
      float e = ez;
      const float2 c = coef<TBL>(tab, g.g1, g.g2, mz);
      if (YEE) e = yee_upd(c.x, e, c.y, curl2(hy0 - hy_im, idxh, hx0 - hx_jm, idyh));
      if (CPML) {
        if (syn) {  // psi_Ezy, left/right (:630-642)
          const long long s = ((long long)k * (g.L[1][0] + g.L[1][1]) + qyn) * (nx + 1) + i;
          e = madd(e, -c.y, psi_step(g.psi[PSI_EZY], s, g.pbe[1][j + 1], hx0 - hx_jm, g.pce[1][j + 1]));
        }
        if (sxn) {  // psi_Ezx, back/front (:771-783)
          const long long s = ((long long)k * (ny + 1) + j) * (g.L[0][0] + g.L[0][1]) + qxn;
          e = madd(e, c.y, psi_step(g.psi[PSI_EZX], s, g.pbe[0][i + 1], hy0 - hy_im, g.pce[0][i + 1]));
        }
      }
      Ez[p] = e;
    }
    if (NEEDH) { hx_km = hx0; hy_km = hy0; }
  }
}

// one tile of the H face pass
template <bool YEE, bool CPML, bool TBL>
__device__ __forceinline__ void hpass_tile(const DevGrid &g, const TileMap &tm,
                                           const MedTab &tab, int blk) {
  int bx, by, bz;
  decode_tile(tm, blk, bx, by, bz);
  const int i = bx * kWave + threadIdx.x;
  const int j = by * blockDim.y + threadIdx.y;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (i > nx || j > ny) return;
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
      const float2 c = coef<TBL>(tab, g.gm1, g.gm2, g.mi[3][p]);
      float h = Hx[p];
      if (YEE) h = yee_upd(c.x, h, c.y, curl2(ey_kp - ey_c, idze, ez_jp - ez_c, idye));
      if (CPML) {
        if (syc) {  // psi_Hxy (borderscpml.F90:911-921)
          const long long s = ((long long)k * (g.L[1][0] + g.L[1][1]) + qyc) * (nx + 1) + i;
          h = madd(h, -c.y, psi_step(g.psi[PSI_HXY], s, g.pbm[1][j + 1], ez_jp - ez_c, g.pcm[1][j + 1]));
        }
        if (szc) {  // psi_Hxz (:1021-1031)
          const long long s = ((long long)qzc * ny + j) * (nx + 1) + i;
          h = madd(h, c.y, psi_step(g.psi[PSI_HXZ], s, g.pbm[2][k + 1], ey_kp - ey_c, g.pcm[2][k + 1]));
        }
      }
      Hx[p] = h;
    }
    if (i < nx && k < nz) {  // Hy sweep: x[0,nx-1] y[0,ny] z[0,nz-1]
      const float2 c = coef<TBL>(tab, g.gm1, g.gm2, g.mi[4][p]);
      float h = Hy[p];
      if (YEE) h = yee_upd(c.x, h, c.y, curl2(ez_ip - ez_c, idxe, ex_kp - ex_c, idze));
      if (CPML) {
        if (szc) {  // psi_Hyz (:1001-1011)
          const long long s = ((long long)qzc * (ny + 1) + j) * nx + i;
          h = madd(h, -c.y, psi_step(g.psi[PSI_HYZ], s, g.pbm[2][k + 1], ex_kp - ex_c, g.pcm[2][k + 1]));
        }
        if (sxc) {  // psi_Hyx (:1111-1121)
          const long long s = ((long long)k * (ny + 1) + j) * (g.L[0][0] + g.L[0][1]) + qxc;
          h = madd(h, c.y, psi_step(g.psi[PSI_HYX], s, g.pbm[0][i + 1], ez_ip - ez_c, g.pcm[0][i + 1]));
        }
      }
      Hy[p] = h;
    }
    if (i < nx && j < ny) {  // Hz sweep: x[0,nx-1] y[0,ny-1] z[0,nz]
      const float2 c = coef<TBL>(tab, g.gm1, g.gm2, g.mi[5][p]);
      float h = Hz[p];
      if (YEE) h = yee_upd(c.x, h, c.y, curl2(ex_jp - ex_c, idye, ey_ip - ey_c, idxe));
      if (CPML) {
        if (syc) {  // psi_Hzy (:931-941)
          const long long s = ((long long)k * (g.L[1][0] + g.L[1][1]) + qyc) * nx + i;
          h = madd(h, c.y, psi_step(g.psi[PSI_HZY], s, g.pbm[1][j + 1], ex_jp - ex_c, g.pcm[1][j + 1]));
        }
        if (sxc) {  // psi_Hzx (:1091-1101)
          const long long s = ((long long)k * ny + j) * (g.L[0][0] + g.L[0][1]) + qxc;
          h = madd(h, -c.y, psi_step(g.psi[PSI_HZX], s, g.pbm[0][i + 1], ey_ip - ey_c, g.pcm[0][i + 1]));
        }
      }
      Hz[p] = h;
    }
    ex_c = ex_kp;
    ey_c = ey_kp;
  }
}

// ---- four consecutive x cells per lane ------------------------------------
// The fused passes move each field value with 16-byte (fp32) accesses: lane l
// of a wave owns cells 4l..4l+3 of a 256-cell x segment.  The x neighbour one
// cell down (E pass) or up (H pass) comes from the next lane by a cross-lane
// shuffle, and from one extra load at the wave's edge.  The row pitch leaves
// room for the last vector of a row, whose cells beyond the sweep are written
// back unchanged.
struct F4 {
  float v[4];
};

template <class T> struct Store;
// fp32 fields: plain 16-byte vectors (the scale arguments are unused)
template <> struct Store<float> {
  static __device__ __forceinline__ F4 ld4(const float *b, long long p, float) {
    const float4 t = *reinterpret_cast<const float4 *>(b + p);
    return {{t.x, t.y, t.z, t.w}};
  }
  static __device__ __forceinline__ float ld1(const float *b, long long p, float) { return b[p]; }
  static __device__ __forceinline__ void st4(float *b, long long p, float, const F4 &v) {
    *reinterpret_cast<float4 *>(b + p) = make_float4(v.v[0], v.v[1], v.v[2], v.v[3]);
  }
};
// binary16 fields with a per-array scale: value = half * s, stored as
// round(value / s); 8-byte vectors, fp32 arithmetic
template <> struct Store<__half> {
  static __device__ __forceinline__ F4 ld4(const float *b, long long p, float s) {
    const uint2 t = *reinterpret_cast<const uint2 *>(reinterpret_cast<const __half *>(b) + p);
    const float2 lo = __half22float2(*reinterpret_cast<const __half2 *>(&t.x));
    const float2 hi = __half22float2(*reinterpret_cast<const __half2 *>(&t.y));
    return {{lo.x * s, lo.y * s, hi.x * s, hi.y * s}};
  }
  static __device__ __forceinline__ float ld1(const float *b, long long p, float s) {
    return __half2float(reinterpret_cast<const __half *>(b)[p]) * s;
  }
  static __device__ __forceinline__ void st4(float *b, long long p, float is, const F4 &v) {
    const __half2 lo = __floats2half2_rn(v.v[0] * is, v.v[1] * is);
    const __half2 hi = __floats2half2_rn(v.v[2] * is, v.v[3] * is);
    uint2 t;
    t.x = *reinterpret_cast<const unsigned *>(&lo);
    t.y = *reinterpret_cast<const unsigned *>(&hi);
    *reinterpret_cast<uint2 *>(reinterpret_cast<__half *>(b) + p) = t;
  }
};

__device__ __forceinline__ void ld_mi4(const int16_t *m, long long p, int (&o)[4]) {
  const short4 t = *reinterpret_cast<const short4 *>(m + p);
  o[0] = t.x; o[1] = t.y; o[2] = t.z; o[3] = t.w;
}

// E face pass (Yee E + CPML E), four cells per lane
template <bool TBL, class T>
__device__ __forceinline__ void epass4_tile(const DevGrid &g, const TileMap &tm,
                                            const MedTab &tab, int blk) {
  int bx, by, bz;
  decode_tile(tm, blk, bx, by, bz);
  const int lane = threadIdx.x;
  const int i0 = bx * (kWave * 4) + 4 * lane;
  const int j = by * blockDim.y + threadIdx.y;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (j > ny) return;                       // whole wave: one row per wave
  const bool live = i0 <= nx;               // lanes past the sweep only shuffle
  const int k0 = bz * tm.kc;
  const int k1 = min(k0 + tm.kc, nz + 1);
  const long long px = g.px, pxy = g.pxy;
  long long p = (long long)(k0 + 1) * pxy + (long long)(j + 1) * px + kXOff + i0;
  float *__restrict__ Ex = g.f[0];
  float *__restrict__ Ey = g.f[1];
  float *__restrict__ Ez = g.f[2];
  const float *__restrict__ Hx = g.f[3];
  const float *__restrict__ Hy = g.f[4];
  const float *__restrict__ Hz = g.f[5];
  F4 hx_km = {}, hy_km = {};
  if (live) {
    hx_km = Store<T>::ld4(Hx, p - pxy, g.fs[3]);
    hy_km = Store<T>::ld4(Hy, p - pxy, g.fs[4]);
  }
  const float idyh = g.id[4][j + 1];
  float idxh[4];
  int qxn[4];
  bool sxn[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int i = min(i0 + c, nx + 1);
    idxh[c] = g.id[3][i + 1];
    sxn[c] = slab_node(i0 + c, nx, g.L[0][0], g.L[0][1], qxn[c]);
  }
  int qyn = 0;
  const bool syn = slab_node(j, ny, g.L[1][0], g.L[1][1], qyn);
  const bool cy = j < ny;
  const int Lx = g.L[0][0] + g.L[0][1], Ly = g.L[1][0] + g.L[1][1];
  for (int k = k0; k < k1; ++k, p += pxy) {
    const bool cz = k < nz;
    const float idzh = g.id[5][k + 1];
    F4 hx0 = {}, hy0 = {}, hz0 = {}, hz_jm = {}, hx_jm = {}, ex = {}, ey = {}, ez = {};
    int mx[4] = {}, my[4] = {}, mz[4] = {};
    float hz_edge = 0.0f, hy_edge = 0.0f;
    if (live) {
      hx0 = Store<T>::ld4(Hx, p, g.fs[3]); hy0 = Store<T>::ld4(Hy, p, g.fs[4]); hz0 = Store<T>::ld4(Hz, p, g.fs[5]);
      hz_jm = Store<T>::ld4(Hz, p - px, g.fs[5]); hx_jm = Store<T>::ld4(Hx, p - px, g.fs[3]);
      ex = Store<T>::ld4(Ex, p, g.fs[0]); ld_mi4(g.mi[0], p, mx);
      if (cy) { ey = Store<T>::ld4(Ey, p, g.fs[1]); ld_mi4(g.mi[1], p, my); }
      if (cz) { ez = Store<T>::ld4(Ez, p, g.fs[2]); ld_mi4(g.mi[2], p, mz); }
      if (lane == 0) { hz_edge = Store<T>::ld1(Hz, p - 1, g.fs[5]); hy_edge = Store<T>::ld1(Hy, p - 1, g.fs[4]); }
    }
    // x - 1 neighbours: the previous lane's last cell
    float hz_im[4], hy_im[4];
    const float hz_up = __shfl_up(hz0.v[3], 1), hy_up = __shfl_up(hy0.v[3], 1);
    hz_im[0] = lane == 0 ? hz_edge : hz_up;
    hy_im[0] = lane == 0 ? hy_edge : hy_up;
#pragma unroll
    for (int c = 1; c < 4; ++c) { hz_im[c] = hz0.v[c - 1]; hy_im[c] = hy0.v[c - 1]; }
    int qzn = 0;
    const bool szn = slab_node(k, nz, g.L[2][0], g.L[2][1], qzn);
    if (live) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int i = i0 + c;
        if (i < nx) {  // Ex sweep x[0,nx-1]
          float e = ex.v[c];
          const float2 cf = coef<TBL>(tab, g.g1, g.g2, mx[c]);
          e = yee_upd(cf.x, e, cf.y, curl2(hz0.v[c] - hz_jm.v[c], idyh, hy0.v[c] - hy_km.v[c], idzh));
          if (syn) {  // psi_Exy, left/right (borderscpml.F90:609-623)
            const long long s = ((long long)k * Ly + qyn) * nx + i;
            e = madd(e, cf.y, psi_step(g.psi[PSI_EXY], s, g.pbe[1][j + 1], hz0.v[c] - hz_jm.v[c],
                                    g.pce[1][j + 1]));
          }
          if (szn) {  // psi_Exz, down/up (:699-712)
            const long long s = ((long long)qzn * (ny + 1) + j) * nx + i;
            e = madd(e, -cf.y, psi_step(g.psi[PSI_EXZ], s, g.pbe[2][k + 1], hy0.v[c] - hy_km.v[c],
                                    g.pce[2][k + 1]));
          }
          ex.v[c] = e;
        }
        if (cy && i <= nx) {  // Ey sweep x[0,nx] y[0,ny-1]
          float e = ey.v[c];
          const float2 cf = coef<TBL>(tab, g.g1, g.g2, my[c]);
          e = yee_upd(cf.x, e, cf.y, curl2(hx0.v[c] - hx_km.v[c], idzh, hz0.v[c] - hz_im[c], idxh[c]));
          if (szn) {  // psi_Eyz, down/up (:680-693)
            const long long s = ((long long)qzn * ny + j) * (nx + 1) + i;
            e = madd(e, cf.y, psi_step(g.psi[PSI_EYZ], s, g.pbe[2][k + 1], hx0.v[c] - hx_km.v[c],
                                    g.pce[2][k + 1]));
          }
          if (sxn[c]) {  // psi_Eyx, back/front (:790-802)
            const long long s = ((long long)k * ny + j) * Lx + qxn[c];
            e = madd(e, -cf.y, psi_step(g.psi[PSI_EYX], s, g.pbe[0][i + 1], hz0.v[c] - hz_im[c],
                                    g.pce[0][i + 1]));
          }
          ey.v[c] = e;
        }
        if (cz && i <= nx) {  // Ez sweep x[0,nx] z[0,nz-1]
          float e = ez.v[c];
          const float2 cf = coef<TBL>(tab, g.g1, g.g2, mz[c]);
          e = yee_upd(cf.x, e, cf.y, curl2(hy0.v[c] - hy_im[c], idxh[c], hx0.v[c] - hx_jm.v[c], idyh));
          if (syn) {  // psi_Ezy, left/right (:630-642)
            const long long s = ((long long)k * Ly + qyn) * (nx + 1) + i;
            e = madd(e, -cf.y, psi_step(g.psi[PSI_EZY], s, g.pbe[1][j + 1], hx0.v[c] - hx_jm.v[c],
                                    g.pce[1][j + 1]));
          }
          if (sxn[c]) {  // psi_Ezx, back/front (:771-783)
            const long long s = ((long long)k * (ny + 1) + j) * Lx + qxn[c];
            e = madd(e, cf.y, psi_step(g.psi[PSI_EZX], s, g.pbe[0][i + 1], hy0.v[c] - hy_im[c],
                                    g.pce[0][i + 1]));
          }
          ez.v[c] = e;
        }
      }
      Store<T>::st4(Ex, p, g.fi[0], ex);
      if (cy) Store<T>::st4(Ey, p, g.fi[1], ey);
      if (cz) Store<T>::st4(Ez, p, g.fi[2], ez);
    }
    hx_km = hx0;
    hy_km = hy0;
  }
}

// H face pass (Yee H + CPML H), four cells per lane
template <bool TBL, class T>
__device__ __forceinline__ void hpass4_tile(const DevGrid &g, const TileMap &tm,
                                            const MedTab &tab, int blk) {
  int bx, by, bz;
  decode_tile(tm, blk, bx, by, bz);
  const int lane = threadIdx.x;
  const int i0 = bx * (kWave * 4) + 4 * lane;
  const int j = by * blockDim.y + threadIdx.y;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (j > ny) return;
  const bool live = i0 <= nx;
  const int k0 = bz * tm.kc;
  const int k1 = min(k0 + tm.kc, nz + 1);
  const long long px = g.px, pxy = g.pxy;
  long long p = (long long)(k0 + 1) * pxy + (long long)(j + 1) * px + kXOff + i0;
  const float *__restrict__ Ex = g.f[0];
  const float *__restrict__ Ey = g.f[1];
  const float *__restrict__ Ez = g.f[2];
  float *__restrict__ Hx = g.f[3];
  float *__restrict__ Hy = g.f[4];
  float *__restrict__ Hz = g.f[5];
  const float idye = g.id[1][j + 1];
  float idxe[4];
  int qxc[4];
  bool sxc[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int i = min(i0 + c, nx + 1);
    idxe[c] = g.id[0][i + 1];
    sxc[c] = slab_cell(i0 + c, nx, g.L[0][0], g.L[0][1], qxc[c]);
  }
  int qyc = 0;
  const bool syc = slab_cell(j, ny, g.L[1][0], g.L[1][1], qyc);
  const bool cy = j < ny;
  const int Lx = g.L[0][0] + g.L[0][1], Ly = g.L[1][0] + g.L[1][1];
  F4 ex_c = {}, ey_c = {};
  if (live) { ex_c = Store<T>::ld4(Ex, p, g.fs[0]); ey_c = Store<T>::ld4(Ey, p, g.fs[1]); }
  for (int k = k0; k < k1; ++k, p += pxy) {
    const float idze = g.id[2][k + 1];
    const bool cz = k < nz;
    F4 ex_kp = {}, ey_kp = {}, ez_c = {}, ez_jp = {}, ex_jp = {}, hx = {}, hy = {}, hz = {};
    int mx[4] = {}, my[4] = {}, mz[4] = {};
    float ez_edge = 0.0f, ey_edge = 0.0f;
    if (live) {
      ex_kp = Store<T>::ld4(Ex, p + pxy, g.fs[0]); ey_kp = Store<T>::ld4(Ey, p + pxy, g.fs[1]);
      ez_c = Store<T>::ld4(Ez, p, g.fs[2]);
      ez_jp = Store<T>::ld4(Ez, p + px, g.fs[2]); ex_jp = Store<T>::ld4(Ex, p + px, g.fs[0]);
      if (cy && cz) { hx = Store<T>::ld4(Hx, p, g.fs[3]); ld_mi4(g.mi[3], p, mx); }
      if (cz) { hy = Store<T>::ld4(Hy, p, g.fs[4]); ld_mi4(g.mi[4], p, my); }
      if (cy) { hz = Store<T>::ld4(Hz, p, g.fs[5]); ld_mi4(g.mi[5], p, mz); }
      if (lane == kWave - 1) { ez_edge = Store<T>::ld1(Ez, p + 4, g.fs[2]); ey_edge = Store<T>::ld1(Ey, p + 4, g.fs[1]); }
    }
    // x + 1 neighbours: the next lane's first cell
    float ez_ip[4], ey_ip[4];
    const float ez_dn = __shfl_down(ez_c.v[0], 1), ey_dn = __shfl_down(ey_c.v[0], 1);
#pragma unroll
    for (int c = 0; c < 3; ++c) { ez_ip[c] = ez_c.v[c + 1]; ey_ip[c] = ey_c.v[c + 1]; }
    ez_ip[3] = lane == kWave - 1 ? ez_edge : ez_dn;
    ey_ip[3] = lane == kWave - 1 ? ey_edge : ey_dn;
    int qzc = 0;
    const bool szc = slab_cell(k, nz, g.L[2][0], g.L[2][1], qzc);
    if (live) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int i = i0 + c;
        if (cy && cz && i <= nx) {  // Hx sweep: x[0,nx] y[0,ny-1] z[0,nz-1]
          const float2 cf = coef<TBL>(tab, g.gm1, g.gm2, mx[c]);
          float h = hx.v[c];
          h = yee_upd(cf.x, h, cf.y, curl2(ey_kp.v[c] - ey_c.v[c], idze, ez_jp.v[c] - ez_c.v[c], idye));
          if (syc) {  // psi_Hxy (borderscpml.F90:911-921)
            const long long s = ((long long)k * Ly + qyc) * (nx + 1) + i;
            h = madd(h, -cf.y, psi_step(g.psi[PSI_HXY], s, g.pbm[1][j + 1], ez_jp.v[c] - ez_c.v[c],
                                    g.pcm[1][j + 1]));
          }
          if (szc) {  // psi_Hxz (:1021-1031)
            const long long s = ((long long)qzc * ny + j) * (nx + 1) + i;
            h = madd(h, cf.y, psi_step(g.psi[PSI_HXZ], s, g.pbm[2][k + 1], ey_kp.v[c] - ey_c.v[c],
                                    g.pcm[2][k + 1]));
          }
          hx.v[c] = h;
        }
        if (cz && i < nx) {  // Hy sweep: x[0,nx-1] y[0,ny] z[0,nz-1]
          const float2 cf = coef<TBL>(tab, g.gm1, g.gm2, my[c]);
          float h = hy.v[c];
          h = yee_upd(cf.x, h, cf.y, curl2(ez_ip[c] - ez_c.v[c], idxe[c], ex_kp.v[c] - ex_c.v[c], idze));
          if (szc) {  // psi_Hyz (:1001-1011)
            const long long s = ((long long)qzc * (ny + 1) + j) * nx + i;
            h = madd(h, -cf.y, psi_step(g.psi[PSI_HYZ], s, g.pbm[2][k + 1], ex_kp.v[c] - ex_c.v[c],
                                    g.pcm[2][k + 1]));
          }
          if (sxc[c]) {  // psi_Hyx (:1111-1121)
            const long long s = ((long long)k * (ny + 1) + j) * Lx + qxc[c];
            h = madd(h, cf.y, psi_step(g.psi[PSI_HYX], s, g.pbm[0][i + 1], ez_ip[c] - ez_c.v[c],
                                    g.pcm[0][i + 1]));
          }
          hy.v[c] = h;
        }
        if (cy && i < nx) {  // Hz sweep: x[0,nx-1] y[0,ny-1] z[0,nz]
          const float2 cf = coef<TBL>(tab, g.gm1, g.gm2, mz[c]);
          float h = hz.v[c];
          h = yee_upd(cf.x, h, cf.y, curl2(ex_jp.v[c] - ex_c.v[c], idye, ey_ip[c] - ey_c.v[c], idxe[c]));
          if (syc) {  // psi_Hzy (:931-941)
            const long long s = ((long long)k * Ly + qyc) * nx + i;
            h = madd(h, cf.y, psi_step(g.psi[PSI_HZY], s, g.pbm[1][j + 1], ex_jp.v[c] - ex_c.v[c],
                                    g.pcm[1][j + 1]));
          }
          if (sxc[c]) {  // psi_Hzx (:1091-1101)
            const long long s = ((long long)k * ny + j) * Lx + qxc[c];
            h = madd(h, -cf.y, psi_step(g.psi[PSI_HZX], s, g.pbm[0][i + 1], ey_ip[c] - ey_c.v[c],
                                    g.pcm[0][i + 1]));
          }
          hz.v[c] = h;
        }
      }
      if (cy && cz) Store<T>::st4(Hx, p, g.fi[3], hx);
      if (cz) Store<T>::st4(Hy, p, g.fi[4], hy);
      if (cy) Store<T>::st4(Hz, p, g.fi[5], hz);
    }
    ex_c = ex_kp;
    ey_c = ey_kp;
  }
}

template <class T>
__global__ void __launch_bounds__(256) k_epass4(DevGrid g, TileMap tm) {
  extern __shared__ float2 lds_tab[];
  const MedTab tab = stage_tables<true, false>(g, g.g1, g.g2, lds_tab);
  epass4_tile<true, T>(g, tm, tab, blockIdx.x);
}

template <class T>
__global__ void __launch_bounds__(256) k_hpass4(DevGrid g, TileMap tm) {
  extern __shared__ float2 lds_tab[];
  const MedTab tab = stage_tables<true, false>(g, g.gm1, g.gm2, lds_tab);
  hpass4_tile<true, T>(g, tm, tab, blockIdx.x);
}

template <bool YEE, bool CPML, bool TBL>
__global__ void __launch_bounds__(256) k_epass(DevGrid g, TileMap tm) {
  extern __shared__ float2 lds_tab[];
  const MedTab tab = stage_tables<TBL, false>(g, g.g1, g.g2, lds_tab);
  epass_tile<YEE, CPML, TBL>(g, tm, tab, blockIdx.x);
}

template <bool YEE, bool CPML, bool TBL>
__global__ void __launch_bounds__(256) k_hpass(DevGrid g, TileMap tm) {
  extern __shared__ float2 lds_tab[];
  const MedTab tab = stage_tables<TBL, false>(g, g.gm1, g.gm2, lds_tab);
  hpass_tile<YEE, CPML, TBL>(g, tm, tab, blockIdx.x);
}

// ---- batches of independent grids ----------------------------------------
// One persistent launch per half step covers every tile of every grid of the
// batch: block b takes tiles b, b + gridDim.x, ...; the grid of a tile is
// found by binary search in the prefix of tile counts, its media tables are
// staged in LDS before the tile runs.
__device__ __forceinline__ int batch_grid(const int *start, int ngrid, int t) {
  int lo = 0, hi = ngrid - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= t) lo = mid; else hi = mid - 1;
  }
  return lo;
}

template <bool H, class T, int VEC>
__global__ void __launch_bounds__(256) k_pass_batch(BatchDev bd) {
  extern __shared__ float2 lds_tab[];
  for (int t = blockIdx.x; t < bd.total; t += gridDim.x) {
    const int b = batch_grid(bd.start, bd.ngrid, t);
    const DevGrid &g = bd.grids[b];
    const TileMap &tm = bd.tms[b];
    __syncthreads();  // the previous tile is done with the LDS tables
    const MedTab tab = H ? stage_tables<true, false>(g, g.gm1, g.gm2, lds_tab)
                         : stage_tables<true, false>(g, g.g1, g.g2, lds_tab);
    // the CPML variant serves grids without CPML too: with no slab layers
    // no cell of the sweep falls in a slab
    const int blk = t - bd.start[b];
    if (VEC == 4) {
      if (H) hpass4_tile<true, T>(g, tm, tab, blk);
      else epass4_tile<true, T>(g, tm, tab, blk);
    } else {
      if (H) hpass_tile<true, true, true>(g, tm, tab, blk);
      else epass_tile<true, true, true>(g, tm, tab, blk);
    }
  }
}

int launch_pass_batch(const BatchDev &bd, int half, int blocks, int rows, int vec, size_t lds,
                      int fp16, void *stream) {
  if (bd.total <= 0) return 0;
  const dim3 blk(kWave, (unsigned)rows);
  hipStream_t s = (hipStream_t)stream;
  if (fp16) {
    if (half) hipLaunchKernelGGL((k_pass_batch<true, __half, 4>), dim3(blocks), blk, lds, s, bd);
    else hipLaunchKernelGGL((k_pass_batch<false, __half, 4>), dim3(blocks), blk, lds, s, bd);
  } else if (vec == 4) {
    if (half) hipLaunchKernelGGL((k_pass_batch<true, float, 4>), dim3(blocks), blk, lds, s, bd);
    else hipLaunchKernelGGL((k_pass_batch<false, float, 4>), dim3(blocks), blk, lds, s, bd);
  } else {
    if (half) hipLaunchKernelGGL((k_pass_batch<true, float, 1>), dim3(blocks), blk, lds, s, bd);
    else hipLaunchKernelGGL((k_pass_batch<false, float, 1>), dim3(blocks), blk, lds, s, bd);
  }
  return (int)hipGetLastError();
}

// media tables fit in LDS up to this many entries
constexpr int kLdsMedia = 4096;

static size_t e_lds_bytes(const DevGrid &g) { return (size_t)(g.nmedia + 1) * 12; }

static size_t ade_lds_bytes(const DevGrid &g, int side) {
  const size_t nd = (size_t)g.ade[side].ndisp;
  return ((e_lds_bytes(g) + nd * 4 + 7) & ~(size_t)7) + nd * kPMax * 6 * sizeof(coef_t);
}

// the pole stage of one field, after its Yee+CPML pass (AdvanceEDispersiveE
// follows AdvanceelectricCPML, timestepping.F90:1380-1417; AdvanceMdispersiveH
// follows the magnetic CPML and PMC clone, :1533-1587)
static void launch_ade(const DevGrid &g, int side, hipStream_t s) {
  const AdeSet &ad = g.ade[side];
  if (!ad.has) return;
  long long nmax = 0;
  for (int c = 0; c < 3; ++c) nmax = nmax > ad.nstate[c] ? nmax : ad.nstate[c];
  if (nmax <= 0) return;
  const size_t la = ade_lds_bytes(g, side);
  const bool ta = g.nmedia + 1 <= kLdsMedia && la <= 48 * 1024;
  const long long want = (nmax + 255) / 256;
  const dim3 grid((unsigned)(want < 8192 ? want : 8192), 3);
  if (side == 0) {
    if (ta) hipLaunchKernelGGL((k_ade_box<true, 0>), grid, dim3(256), la, s, g);
    else hipLaunchKernelGGL((k_ade_box<false, 0>), grid, dim3(256), 0, s, g);
  } else {
    if (ta) hipLaunchKernelGGL((k_ade_box<true, 1>), grid, dim3(256), la, s, g);
    else hipLaunchKernelGGL((k_ade_box<false, 1>), grid, dim3(256), 0, s, g);
  }
}

int launch_e(const DevGrid &g, const TileMap &tm, const TileMap &tm4, int yee, int cpml,
             int ade, void *stream) {
  const dim3 blk(kWave, (unsigned)tm.bdy);
  hipStream_t s = (hipStream_t)stream;
  const int nb = tm.nblk;
  const size_t lds = e_lds_bytes(g);
  const bool tbl = g.nmedia + 1 <= kLdsMedia;
#define SFD_E(Y, C)                                                            \
  do {                                                                         \
    if (tbl)                                                                   \
      hipLaunchKernelGGL((k_epass<Y, C, true>), dim3(nb), blk, lds, s, g, tm);  \
    else                                                                       \
      hipLaunchKernelGGL((k_epass<Y, C, false>), dim3(nb), blk, 0, s, g, tm);   \
  } while (0)
  if (yee && cpml && tbl && (tm4.vec == 4 || g.fp16)) {
    // fused pass, four cells per lane (the CPML variant also serves grids
    // without CPML)
    if (g.fp16) hipLaunchKernelGGL(k_epass4<__half>, dim3(tm4.nblk), blk, lds, s, g, tm4);
    else hipLaunchKernelGGL(k_epass4<float>, dim3(tm4.nblk), blk, lds, s, g, tm4);
  } else if (yee && cpml) SFD_E(true, true);
  else if (yee) SFD_E(true, false);
  else if (cpml) SFD_E(false, true);
#undef SFD_E
  if (ade) launch_ade(g, 0, s);
  return (int)hipGetLastError();
}

int launch_h(const DevGrid &g, const TileMap &tm, const TileMap &tm4, int yee, int cpml,
             int ade, void *stream) {
  const dim3 blk(kWave, (unsigned)tm.bdy);
  hipStream_t s = (hipStream_t)stream;
  const int nb = tm.nblk;
  const bool tbl = g.nmedia + 1 <= kLdsMedia;
  const size_t lds = tbl ? (size_t)(g.nmedia + 1) * sizeof(float2) : 0;
#define SFD_H(Y, C)                                                            \
  do {                                                                         \
    if (tbl)                                                                   \
      hipLaunchKernelGGL((k_hpass<Y, C, true>), dim3(nb), blk, lds, s, g, tm);  \
    else                                                                       \
      hipLaunchKernelGGL((k_hpass<Y, C, false>), dim3(nb), blk, 0, s, g, tm);   \
  } while (0)
  if (yee && cpml && tbl && (tm4.vec == 4 || g.fp16)) {
    if (g.fp16) hipLaunchKernelGGL(k_hpass4<__half>, dim3(tm4.nblk), blk, lds, s, g, tm4);
    else hipLaunchKernelGGL(k_hpass4<float>, dim3(tm4.nblk), blk, lds, s, g, tm4);
  } else if (yee && cpml) SFD_H(true, true);
  else if (yee) SFD_H(true, false);
  else if (cpml) SFD_H(false, true);
#undef SFD_H
  if (ade) launch_ade(g, 1, s);
  return (int)hipGetLastError();
}

// ---- one launch per step ---------------------------------------------------
// k_step advances every field of the grid by one step (E to n+1, then H to
// n+1/2), reading the state of step n from sd.gi and writing step n+1 to
// sd.go: fields, CPML psi and pole state all live twice, and the two sets
// alternate.  Per cell the arithmetic and its order are those of the two
// passes (Yee, psi corrections, poles, plane-wave faces, nodal cells), so the
// result is the same bit for bit; only the traffic changes: each field value
// is read once and written once per step instead of twice.
//
// A block of 64 x R threads marches one z chunk of a tile.  Lanes 0..62 and
// rows 0..R-2 own their cells; lane 63 and row R-1 overlap the neighbouring
// tiles and compute only the E values the owned H cells need, and the plane
// above the chunk is computed the same way.  E values travel to the x - 1
// neighbour by a cross-lane shuffle and to the y - 1 neighbour through LDS;
// H values of step n are read from memory (gi is never written during the
// step).

// psi <- b*psi + d*c from gi's array, written to go's array when wr
__device__ __forceinline__ float psi_next(const float *pi, float *po, bool wr, long long s,
                                          float b, float d, float c) {
  const float v = madd(d * c, b, pi[s]);
  if (wr) po[s] = v;
  return v;
}

// the pole state of one dispersive edge of component C (side: E or H)
template <int C>
__device__ __forceinline__ float step_poles(const AdeSet &ai, const AdeSet &ao, bool wr,
                                            double dt, const MedTab &t, int slot, int i, int j,
                                            int k, float e) {
  const int *b = ai.abox[C];
  AdePre a;
  a.s = ((long long)(k - b[2]) * b[4] + (j - b[1])) * b[3] + (i - b[0]);
  a.slot = slot;
  a.np = t.anp[slot];
  a.ep = ai.prev[C][a.s];
  const jstate_t *J = reinterpret_cast<const jstate_t *>(ai.jst[C]);
#pragma unroll
  for (int q = 0; q < kPre; ++q)
    if (q < a.np) a.j[q] = J[q * ai.nstate[C] + a.s];
  return ade_apply<C>(ai, ao, wr, dt, t, a, e);
}

// plane-wave faces and nodal cells of one half acting on component comp at
// cell cc (offset p), in the reference's order
__device__ __noinline__ float step_sources(const StepDev &sd, int half, int comp, long long p,
                                           int i, int j, int k, float v, bool pw, bool nd,
                                           int n) {
  if (pw) {
    const int cc[3] = {i, j, k};
    const float tm = half ? (float)(sd.tiempo[n] + (double)0.5f * sd.dt) : (float)sd.tiempo[n];
    for (int q = 0; q < sd.npw[half]; ++q) {
      const PwFace &f = sd.pwf[half][q];
      if (f.comp != comp || cc[f.axis] != f.fixed) continue;
      const int ua = f.axis == 0 ? 1 : 0, va = f.axis == 2 ? 1 : 2;
      if (cc[ua] < f.lo[ua] || cc[ua] > f.hi[ua] || cc[va] < f.lo[va] || cc[va] > f.hi[va])
        continue;
      const float ehi = pw_incid(f, sd.waves, sd.phys, sd.pstride, tm, cc);
      v = pw_apply(v, f.sign, sd.pwG[half], ehi, f.id);
    }
  }
  if (nd) {
    for (int q = 0; q < sd.ncell[half]; ++q) {
      const NodalCell &c = sd.cells[half][q];
      if (c.comp != comp || c.p != p || c.skip) continue;
      v = nodal_apply(v, c, sd.values[(long long)n * sd.nsrc + c.src]);
    }
  }
  return v;
}

// does a wave of row j, plane k, columns [x0, x0+63] touch a TF/SF face of
// the half: the x and y tests do not change along the z march
__device__ __forceinline__ bool step_pw_xy(const StepDev &sd, int half, int x0, int j) {
  if (sd.npw[half] == 0) return false;
  if (sd.npl[half][0] < 0 || sd.npl[half][1] < 0 || sd.npl[half][2] < 0) return true;
  for (int q = 0; q < sd.npl[half][0]; ++q)
    if (sd.pl[half][0][q] >= x0 && sd.pl[half][0][q] <= x0 + kWave - 1) return true;
  for (int q = 0; q < sd.npl[half][1]; ++q)
    if (sd.pl[half][1][q] == j) return true;
  return false;
}
__device__ __forceinline__ bool step_pw_z(const StepDev &sd, int half, int k) {
  for (int q = 0; q < sd.npl[half][2]; ++q)
    if (sd.pl[half][2][q] == k) return true;
  return false;
}
__device__ __forceinline__ bool step_nd(const StepDev &sd, int half, int x0, int j, int k) {
  const int *b = sd.cellbox[half];
  return sd.ncell[half] > 0 && j >= b[1] && j <= b[4] && k >= b[2] && k <= b[5] &&
         x0 + kWave - 1 >= b[0] && x0 <= b[3];
}

static __host__ __device__ __forceinline__ size_t step_tab_bytes(int nmedia, int nd) {
  return (((size_t)(nmedia + 1) * 12 + (size_t)nd * 4 + 7) & ~(size_t)7) +
         (size_t)nd * kPMax * 6 * sizeof(coef_t);
}

// ADE: grids with pole state on either field; SRC: grids with plane waves or
// nodal sources (the variants without them carry no code for them)
template <int R, bool ADE, bool SRC>
__global__ void __launch_bounds__(64 * R) k_step(StepDev sd) {
  extern __shared__ float2 lds[];
  const DevGrid &gi = sd.gi;
  const DevGrid &go = sd.go;
  const size_t offh = step_tab_bytes(gi.nmedia, gi.ade[0].ndisp);
  const size_t offx = offh + step_tab_bytes(gi.nmedia, gi.ade[1].ndisp);
  const MedTab te = stage_tables<true, true>(gi, gi.g1, gi.g2, lds, &gi.ade[0]);
  const MedTab th = stage_tables<true, true>(
      gi, gi.gm1, gi.gm2, reinterpret_cast<float2 *>(reinterpret_cast<char *>(lds) + offh),
      &gi.ade[1]);
  float2 *xch = reinterpret_cast<float2 *>(reinterpret_cast<char *>(lds) + offx);

  int bx, by, bz;
  decode_tile(sd.tm, blockIdx.x, bx, by, bz);
  const int lane = threadIdx.x, row = threadIdx.y;
  const int x0 = bx * (kWave - 1);
  const int i = x0 + lane, j = by * (R - 1) + row;
  const int nx = gi.nx, ny = gi.ny, nz = gi.nz;
  const bool in = i <= nx && j <= ny;
  const bool own = in && lane < kWave - 1 && row < R - 1;
  // threads past the grid's edge read the edge cell (their results are
  // never used): every load and table index stays inside its array
  const int ic = min(i, nx), jc = min(j, ny);
  const int k0 = bz * sd.tm.kc;
  const int k1 = min(k0 + sd.tm.kc, nz + 1);
  const long long px = gi.px, pxy = gi.pxy;
  long long p = (long long)(k0 + 1) * pxy + (long long)(jc + 1) * px + kXOff + ic;
  const int n = *sd.dn;
  const double dt = gi.dt;
  const int Lx = gi.L[0][0] + gi.L[0][1], Ly = gi.L[1][0] + gi.L[1][1];
  // E side: node positions along x and y; H side: cell positions
  int qxn = 0, qyn = 0, qxc = 0, qyc = 0;
  const bool sxn = slab_node(i, nx, gi.L[0][0], gi.L[0][1], qxn);
  const bool syn = slab_node(j, ny, gi.L[1][0], gi.L[1][1], qyn);
  const bool sxc = slab_cell(i, nx, gi.L[0][0], gi.L[0][1], qxc);
  const bool syc = slab_cell(j, ny, gi.L[1][0], gi.L[1][1], qyc);
  const float idyh = gi.id[4][jc + 1], idxh = gi.id[3][ic + 1];
  const float idye = gi.id[1][jc + 1], idxe = gi.id[0][ic + 1];
  const bool pwxy_e = SRC && step_pw_xy(sd, 0, x0, j), pwxy_h = SRC && step_pw_xy(sd, 1, x0, j);
  const bool ade_e = ADE && gi.ade[0].has, ade_h = ADE && gi.ade[1].has;

  // H of step n on plane k0 - 1: the E update's z - 1 neighbours
  float hx_km = gi.f[3][p - pxy], hy_km = gi.f[4][p - pxy];
  // carried from plane k - 1 to the H update of that plane
  float hx_c = 0.0f, hy_c = 0.0f, hz_c = 0.0f;          // H of step n
  float ex_c = 0.0f, ey_c = 0.0f, ez_c = 0.0f;          // E of step n+1
  float ey_ip_c = 0.0f, ez_ip_c = 0.0f, ex_jp_c = 0.0f, ez_jp_c = 0.0f;
  int mhx_c = 0, mhy_c = 0, mhz_c = 0;

  for (int k = k0; k <= k1; ++k, p += pxy) {
    float ex_n = 0.0f, ey_n = 0.0f, ez_n = 0.0f;
    float hx0 = 0.0f, hy0 = 0.0f, hz0 = 0.0f;
    int mhx = 0, mhy = 0, mhz = 0;
    if (k <= nz) {
      // ---- issue the loads of plane k
      hx0 = gi.f[3][p]; hy0 = gi.f[4][p]; hz0 = gi.f[5][p];
      const float hz_jm = gi.f[5][p - px], hx_jm = gi.f[3][p - px];
      const float hz_im = gi.f[5][p - 1], hy_im = gi.f[4][p - 1];
      ex_n = gi.f[0][p]; ey_n = gi.f[1][p]; ez_n = gi.f[2][p];
      const int mx = gi.mi[0][p], my = gi.mi[1][p], mz = gi.mi[2][p];
      mhx = gi.mi[3][p]; mhy = gi.mi[4][p]; mhz = gi.mi[5][p];
      const float idzh = gi.id[5][k + 1];
      int qzn = 0;
      const bool szn = slab_node(k, nz, gi.L[2][0], gi.L[2][1], qzn);
      const bool wr = own && k < k1;
      const bool pw = SRC && (pwxy_e || step_pw_z(sd, 0, k));
      const bool nd = SRC && step_nd(sd, 0, x0, j, k);
      // ---- E of step n+1 (Advance_Ex/Ey/Ez, AdvanceelectricCPML,
      // AdvanceEDispersiveE, AdvancePlaneWaveE, AdvanceNodalE)
      if (in && i < nx) {
        const float2 c = te.gg[mx];
        float e = yee_upd(c.x, ex_n, c.y, curl2(hz0 - hz_jm, idyh, hy0 - hy_km, idzh));
        if (syn)
          e = madd(e, c.y, psi_next(gi.psi[PSI_EXY], go.psi[PSI_EXY], wr,
                                    ((long long)k * Ly + qyn) * nx + i, gi.pbe[1][j + 1],
                                    hz0 - hz_jm, gi.pce[1][j + 1]));
        if (szn)
          e = madd(e, -c.y, psi_next(gi.psi[PSI_EXZ], go.psi[PSI_EXZ], wr,
                                     ((long long)qzn * (ny + 1) + j) * nx + i, gi.pbe[2][k + 1],
                                     hy0 - hy_km, gi.pce[2][k + 1]));
        if (ade_e) {
          const int sl = te.slot[mx];
          if (sl >= 0) e = step_poles<0>(gi.ade[0], go.ade[0], wr, dt, te, sl, i, j, k, e);
        }
        if (pw || nd) e = step_sources(sd, 0, 0, p, i, j, k, e, pw, nd, n);
        ex_n = e;
        if (wr) go.f[0][p] = e;
      }
      if (in && j < ny) {
        const float2 c = te.gg[my];
        float e = yee_upd(c.x, ey_n, c.y, curl2(hx0 - hx_km, idzh, hz0 - hz_im, idxh));
        if (szn)
          e = madd(e, c.y, psi_next(gi.psi[PSI_EYZ], go.psi[PSI_EYZ], wr,
                                    ((long long)qzn * ny + j) * (nx + 1) + i, gi.pbe[2][k + 1],
                                    hx0 - hx_km, gi.pce[2][k + 1]));
        if (sxn)
          e = madd(e, -c.y, psi_next(gi.psi[PSI_EYX], go.psi[PSI_EYX], wr,
                                     ((long long)k * ny + j) * Lx + qxn, gi.pbe[0][i + 1],
                                     hz0 - hz_im, gi.pce[0][i + 1]));
        if (ade_e) {
          const int sl = te.slot[my];
          if (sl >= 0) e = step_poles<1>(gi.ade[0], go.ade[0], wr, dt, te, sl, i, j, k, e);
        }
        if (pw || nd) e = step_sources(sd, 0, 1, p, i, j, k, e, pw, nd, n);
        ey_n = e;
        if (wr) go.f[1][p] = e;
      }
      if (in && k < nz) {
        const float2 c = te.gg[mz];
        float e = yee_upd(c.x, ez_n, c.y, curl2(hy0 - hy_im, idxh, hx0 - hx_jm, idyh));
        if (syn)
          e = madd(e, -c.y, psi_next(gi.psi[PSI_EZY], go.psi[PSI_EZY], wr,
                                     ((long long)k * Ly + qyn) * (nx + 1) + i, gi.pbe[1][j + 1],
                                     hx0 - hx_jm, gi.pce[1][j + 1]));
        if (sxn)
          e = madd(e, c.y, psi_next(gi.psi[PSI_EZX], go.psi[PSI_EZX], wr,
                                    ((long long)k * (ny + 1) + j) * Lx + qxn, gi.pbe[0][i + 1],
                                    hy0 - hy_im, gi.pce[0][i + 1]));
        if (ade_e) {
          const int sl = te.slot[mz];
          if (sl >= 0) e = step_poles<2>(gi.ade[0], go.ade[0], wr, dt, te, sl, i, j, k, e);
        }
        if (pw || nd) e = step_sources(sd, 0, 2, p, i, j, k, e, pw, nd, n);
        ez_n = e;
        if (wr) go.f[2][p] = e;
      }
    }
    // ---- E of plane k at x + 1 (next lane) and y + 1 (next row)
    const float ey_ip_n = __shfl_down(ey_n, 1), ez_ip_n = __shfl_down(ez_n, 1);
    const int par = (k - k0) & 1;
    xch[(par * R + row) * kWave + lane] = make_float2(ex_n, ez_n);
    __syncthreads();
    float ex_jp_n = 0.0f, ez_jp_n = 0.0f;
    if (row + 1 < R) {
      const float2 v = xch[(par * R + row + 1) * kWave + lane];
      ex_jp_n = v.x;
      ez_jp_n = v.y;
    }
    // ---- H of step n+1/2 on plane k - 1 (Advance_Hx/Hy/Hz,
    // AdvanceMagneticCPML, AdvanceMdispersiveH, AdvancePlaneWaveH,
    // AdvanceNodalH)
    if (k > k0 && own) {
      const int kh = k - 1;
      const long long ph = p - pxy;
      const float idze = gi.id[2][kh + 1];
      int qzc = 0;
      const bool szc = slab_cell(kh, nz, gi.L[2][0], gi.L[2][1], qzc);
      const bool pw = SRC && (pwxy_h || step_pw_z(sd, 1, kh));
      const bool nd = SRC && step_nd(sd, 1, x0, j, kh);
      if (j < ny && kh < nz) {
        const float2 c = th.gg[mhx_c];
        float h = yee_upd(c.x, hx_c, c.y, curl2(ey_n - ey_c, idze, ez_jp_c - ez_c, idye));
        if (syc)
          h = madd(h, -c.y, psi_next(gi.psi[PSI_HXY], go.psi[PSI_HXY], true,
                                     ((long long)kh * Ly + qyc) * (nx + 1) + i,
                                     gi.pbm[1][j + 1], ez_jp_c - ez_c, gi.pcm[1][j + 1]));
        if (szc)
          h = madd(h, c.y, psi_next(gi.psi[PSI_HXZ], go.psi[PSI_HXZ], true,
                                    ((long long)qzc * ny + j) * (nx + 1) + i, gi.pbm[2][kh + 1],
                                    ey_n - ey_c, gi.pcm[2][kh + 1]));
        if (ade_h) {
          const int sl = th.slot[mhx_c];
          if (sl >= 0) h = step_poles<0>(gi.ade[1], go.ade[1], true, dt, th, sl, i, j, kh, h);
        }
        if (pw || nd) h = step_sources(sd, 1, 3, ph, i, j, kh, h, pw, nd, n);
        go.f[3][ph] = h;
      }
      if (i < nx && kh < nz) {
        const float2 c = th.gg[mhy_c];
        float h = yee_upd(c.x, hy_c, c.y, curl2(ez_ip_c - ez_c, idxe, ex_n - ex_c, idze));
        if (szc)
          h = madd(h, -c.y, psi_next(gi.psi[PSI_HYZ], go.psi[PSI_HYZ], true,
                                     ((long long)qzc * (ny + 1) + j) * nx + i,
                                     gi.pbm[2][kh + 1], ex_n - ex_c, gi.pcm[2][kh + 1]));
        if (sxc)
          h = madd(h, c.y, psi_next(gi.psi[PSI_HYX], go.psi[PSI_HYX], true,
                                    ((long long)kh * (ny + 1) + j) * Lx + qxc, gi.pbm[0][i + 1],
                                    ez_ip_c - ez_c, gi.pcm[0][i + 1]));
        if (ade_h) {
          const int sl = th.slot[mhy_c];
          if (sl >= 0) h = step_poles<1>(gi.ade[1], go.ade[1], true, dt, th, sl, i, j, kh, h);
        }
        if (pw || nd) h = step_sources(sd, 1, 4, ph, i, j, kh, h, pw, nd, n);
        go.f[4][ph] = h;
      }
      if (i < nx && j < ny) {
        const float2 c = th.gg[mhz_c];
        float h = yee_upd(c.x, hz_c, c.y, curl2(ex_jp_c - ex_c, idye, ey_ip_c - ey_c, idxe));
        if (syc)
          h = madd(h, c.y, psi_next(gi.psi[PSI_HZY], go.psi[PSI_HZY], true,
                                    ((long long)kh * Ly + qyc) * nx + i, gi.pbm[1][j + 1],
                                    ex_jp_c - ex_c, gi.pcm[1][j + 1]));
        if (sxc)
          h = madd(h, -c.y, psi_next(gi.psi[PSI_HZX], go.psi[PSI_HZX], true,
                                     ((long long)kh * ny + j) * Lx + qxc, gi.pbm[0][i + 1],
                                     ey_ip_c - ey_c, gi.pcm[0][i + 1]));
        if (ade_h) {
          const int sl = th.slot[mhz_c];
          if (sl >= 0) h = step_poles<2>(gi.ade[1], go.ade[1], true, dt, th, sl, i, j, kh, h);
        }
        if (pw || nd) h = step_sources(sd, 1, 5, ph, i, j, kh, h, pw, nd, n);
        go.f[5][ph] = h;
      }
    }
    // ---- carry plane k to the next iteration
    hx_km = hx0; hy_km = hy0;
    hx_c = hx0; hy_c = hy0; hz_c = hz0;
    ex_c = ex_n; ey_c = ey_n; ez_c = ez_n;
    ey_ip_c = ey_ip_n; ez_ip_c = ez_ip_n; ex_jp_c = ex_jp_n; ez_jp_c = ez_jp_n;
    mhx_c = mhx; mhy_c = mhy; mhz_c = mhz;
  }
}

size_t step_lds_bytes(const DevGrid &g, int rows) {
  return step_tab_bytes(g.nmedia, g.ade[0].ndisp) + step_tab_bytes(g.nmedia, g.ade[1].ndisp) +
         (size_t)2 * rows * kWave * sizeof(float2);
}

int launch_step(const StepDev &sd, int rows, void *stream) {
  const size_t lds = step_lds_bytes(sd.gi, rows);
  hipStream_t s = (hipStream_t)stream;
  const dim3 grid((unsigned)sd.tm.nblk);
  const bool ade = sd.gi.ade[0].has || sd.gi.ade[1].has;
  const bool src = sd.npw[0] || sd.npw[1] || sd.ncell[0] || sd.ncell[1];
#define SFD_STEP(RR, A, S) \
  hipLaunchKernelGGL((k_step<RR, A, S>), grid, dim3(kWave, RR), lds, s, sd)
#define SFD_STEP_R(RR)                     \
  do {                                     \
    if (ade && src) SFD_STEP(RR, true, true);   \
    else if (ade) SFD_STEP(RR, true, false);    \
    else if (src) SFD_STEP(RR, false, true);    \
    else SFD_STEP(RR, false, false);            \
  } while (0)
  if (rows == 4) SFD_STEP_R(4);
  else if (rows == 8) SFD_STEP_R(8);
  else return (int)hipErrorInvalidValue;
#undef SFD_STEP_R
#undef SFD_STEP
  return (int)hipGetLastError();
}

}  // namespace SFD_NS
}  // namespace sfd
