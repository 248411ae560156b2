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

// hardware dispatches block b to XCD b % 8; give each XCD one contiguous
// range of tiles so that x/y neighbours share that XCD's L2
// global memory, spelled out for the pointers the compiler cannot infer it
// for (those read out of memory rather than passed as kernel arguments)
#define SFD_GLOBAL __attribute__((address_space(1)))

__device__ __forceinline__ int xcd_remap(int b, int nblk) {
  const int q = nblk >> 3, r = nblk & 7;
  const int xcd = b & 7, loc = b >> 3;
  return (xcd < r) ? xcd * (q + 1) + loc : r * (q + 1) + (xcd - r) * q + loc;
}

__device__ __forceinline__ void decode_tile(const TileMap &tm, int b, int &bx, int &by,
                                            int &bz) {
  if (tm.remap) b = xcd_remap(b, tm.nblk);
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
// Fprev = F  (F = E with G3, or H with GM3).  cf: the slot's coefficients,
// ep: Fprev; jget(q) gives the state of pole q and jput(q, v) stores its new
// state.  Returns the new F.  FAST arithmetic is written out as fused
// multiply-adds so that every kernel that inlines it rounds the same way.
template <class JGet, class JPut>
__device__ __forceinline__ float ade_core(const coef_t *cf, int np, float ep, double dt,
                                          float e, JGet jget, JPut jput) {
#if SFD_EXACT
  // one pole at a time, rounding E after each (electricdispersive.F90:319-321)
#pragma unroll
  for (int q = 0; q < kPMax; ++q) {
    if (q >= np) break;
    const double2 jv = jget(q);
    const double re = cf[q * 6 + 4] * jv.x - cf[q * 6 + 5] * jv.y;
    e = (float)((double)e - re);
  }
  const float diff = e - ep;
#pragma unroll
  for (int q = 0; q < kPMax; ++q) {
    if (q >= np) break;
    const double2 jv = jget(q);
    const double kr = cf[q * 6 + 0], ki = cf[q * 6 + 1];
    const double kjr = kr * jv.x - ki * jv.y;
    const double kji = kr * jv.y + ki * jv.x;
    const double br = cf[q * 6 + 2] * (double)diff / dt;
    const double bi = cf[q * 6 + 3] * (double)diff / dt;
    jput(q, make_double2(kjr + br, kji + bi));
  }
#else
  // folded fp32 form: beta is pre-divided by dt, G3 of a folded pair doubled
  (void)dt;
  float sum = 0.0f;
#pragma unroll
  for (int q = 0; q < kPMax; ++q) {
    if (q >= np) break;
    const float2 jv = jget(q);
    sum = sum + __builtin_fmaf(cf[q * 6 + 4], jv.x, -(cf[q * 6 + 5] * jv.y));
  }
  e = e - sum;
  const float diff = e - ep;
#pragma unroll
  for (int q = 0; q < kPMax; ++q) {
    if (q >= np) break;
    const float2 jv = jget(q);
    const float kr = cf[q * 6 + 0], ki = cf[q * 6 + 1];
    const float nr = __builtin_fmaf(kr, jv.x, __builtin_fmaf(-ki, jv.y, cf[q * 6 + 2] * diff));
    const float ni = __builtin_fmaf(kr, jv.y, __builtin_fmaf(ki, jv.x, cf[q * 6 + 3] * diff));
    jput(q, make_float2(nr, ni));
  }
#endif
  return e;
}

// the recursion on edge a.s of component C: state read from ai, the new state
// written to ao when wr (the same set for in-place updates)
template <int C>
__device__ __forceinline__ float ade_apply(const AdeSet &ai, const AdeSet &ao, bool wr,
                                           double dt, const MedTab &t, const AdePre &a,
                                           float e) {
  const long long N = ai.nstate[C];
  const jstate_t *Ji = reinterpret_cast<const jstate_t *>(ai.jst[C]) + a.s;
  jstate_t *Jo = reinterpret_cast<jstate_t *>(ao.jst[C]) + a.s;
  e = ade_core(t.acf + (size_t)a.slot * kPMax * 6, a.np, a.ep, dt, e,
               [&](int q) { return q < kPre ? a.j[q] : Ji[q * N]; },
               [&](int q, jstate_t v) {
                 if (wr) Jo[q * N] = v;
               });
  if (wr) ao.prev[C][a.s] = e;
  return e;
}

// The dispersive-edge pass: AdvanceEDispersiveE (SIDE 0) or
// AdvanceMdispersiveH (SIDE 1) over the pole-state box of each component
// (blockIdx.y), after the Yee+CPML pass of that field.  Edges of the box are
// visited x-fastest, so the media index, field, previous-field and state
// loads of a wave are each one contiguous run; non-dispersive edges cost only
// their int16 media index.
// the pole stage of one grid whose descriptor lives in device memory
template <bool TBL, int SIDE>
__device__ __forceinline__ void ade_box_run(const DevGrid &g) {
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
                        g.xoff + b[0] + x;
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

// (the same body as ade_box_run, written out: a reference to the kernel's
// by-value argument would copy the whole descriptor to scratch memory)
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
                        g.xoff + b[0] + x;
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

// the pole stage of several grids of a batch in one launch: blockIdx.z picks
// the grid (list[z], an index into grids), the rest as k_ade_box
template <bool TBL, int SIDE>
__global__ void __launch_bounds__(256) k_ade_many(const DevGrid *grids, const int *list) {
  ade_box_run<TBL, SIDE>(grids[list[blockIdx.z]]);
}

// Loads and stores of the march address a plane through a base pointer that
// is the same for every lane (scalar registers) plus the lane's byte offset
// inside the plane (one 32-bit register).  The plane base is made opaque to
// the compiler, which otherwise hoists a 64-bit (array base + offset) pair per
// array and lane out of the march; it stays in the global address space so
// that the accesses are global, not flat, loads and stores.
template <class T>
__device__ __forceinline__ SFD_GLOBAL T *plane_base(T *base, long long P) {
  SFD_GLOBAL T *q = (SFD_GLOBAL T *)(base + P);
  asm volatile("" : "+s"(q));
  return q;
}
template <class T>
__device__ __forceinline__ T ld_at(const SFD_GLOBAL T *base, unsigned ob) {
  return *reinterpret_cast<const SFD_GLOBAL T *>(reinterpret_cast<const SFD_GLOBAL char *>(base) +
                                                 ob);
}
template <class T>
__device__ __forceinline__ void st_at(SFD_GLOBAL T *base, unsigned ob, T v) {
  *reinterpret_cast<SFD_GLOBAL T *>(reinterpret_cast<SFD_GLOBAL char *>(base) + ob) = v;
}

// an element of a table the kernel never writes, at an index the same for
// every lane: read through the constant address space, i.e. by a scalar load
// that does not join the vector-memory counter shared with the stores
__device__ __forceinline__ float ro_uniform(const float *t, int idx) {
  return ((const __attribute__((address_space(4))) float *)t)[idx];
}

// Field storage of the face passes: each field array holds fp32 values, or
// binary16 values with a per-array scale.
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
  static __device__ __forceinline__ void st1(float *b, long long p, float, float v) { b[p] = v; }
  static __device__ __forceinline__ float ld1(const SFD_GLOBAL float *b, long long p, float) {
    return b[p];
  }
  static __device__ __forceinline__ void st1(SFD_GLOBAL float *b, long long p, float, float v) {
    b[p] = v;
  }
  static __device__ __forceinline__ void st4(float *b, long long p, float, const F4 &v) {
    *reinterpret_cast<float4 *>(b + p) = make_float4(v.v[0], v.v[1], v.v[2], v.v[3]);
  }
  // a plane base and the lane's byte offset in the plane; ldr/cvt split the
  // load from the conversion so that a plane's loads all issue first
  typedef float raw_t;
  static __device__ __forceinline__ raw_t ldr(const SFD_GLOBAL char *b, unsigned ob) {
    return ld_at((const SFD_GLOBAL float *)b, ob);
  }
  static __device__ __forceinline__ float cvt(raw_t r, float) { return r; }
  static __device__ __forceinline__ float ldo(const SFD_GLOBAL char *b, unsigned ob, float) {
    return ld_at((const SFD_GLOBAL float *)b, ob);
  }
  static __device__ __forceinline__ void sto(SFD_GLOBAL char *b, unsigned ob, float, float v) {
    st_at((SFD_GLOBAL float *)b, ob, v);
  }
};
// binary16 fields with a per-array scale: value = half * s, stored as
// round(value / s); 8-byte vectors, fp32 arithmetic.  Every scale product is
// rounded to fp32 on its own (an empty asm barrier keeps the compiler from
// fusing it into the arithmetic that uses it, or into the conversion to
// binary16 as a mixed-precision FMA, which would round once instead of
// twice), so every kernel shape rounds alike.
__device__ __forceinline__ float rounded(float x) {
  asm("" : "+v"(x));
  return x;
}
template <> struct Store<__half> {
  static __device__ __forceinline__ F4 ld4(const float *b, long long p, float s) {
    const uint2 t = *reinterpret_cast<const uint2 *>(reinterpret_cast<const __half *>(b) + p);
    const float2 lo = __half22float2(*reinterpret_cast<const __half2 *>(&t.x));
    const float2 hi = __half22float2(*reinterpret_cast<const __half2 *>(&t.y));
    return {{rounded(lo.x * s), rounded(lo.y * s), rounded(hi.x * s), rounded(hi.y * s)}};
  }
  static __device__ __forceinline__ float ld1(const float *b, long long p, float s) {
    return rounded(__half2float(reinterpret_cast<const __half *>(b)[p]) * s);
  }
  static __device__ __forceinline__ void st1(float *b, long long p, float is, float v) {
    reinterpret_cast<__half *>(b)[p] = __float2half_rn(rounded(v * is));
  }
  static __device__ __forceinline__ float ld1(const SFD_GLOBAL float *b, long long p, float s) {
    const unsigned short h = ((const SFD_GLOBAL unsigned short *)b)[p];
    return rounded(__half2float(__ushort_as_half(h)) * s);
  }
  static __device__ __forceinline__ void st1(SFD_GLOBAL float *b, long long p, float is, float v) {
    ((SFD_GLOBAL unsigned short *)b)[p] = __half_as_ushort(__float2half_rn(rounded(v * is)));
  }
  static __device__ __forceinline__ void st4(float *b, long long p, float is, const F4 &v) {
    const __half2 lo = __floats2half2_rn(rounded(v.v[0] * is), rounded(v.v[1] * is));
    const __half2 hi = __floats2half2_rn(rounded(v.v[2] * is), rounded(v.v[3] * is));
    uint2 t;
    t.x = *reinterpret_cast<const unsigned *>(&lo);
    t.y = *reinterpret_cast<const unsigned *>(&hi);
    *reinterpret_cast<uint2 *>(reinterpret_cast<__half *>(b) + p) = t;
  }
  typedef unsigned short raw_t;
  static __device__ __forceinline__ raw_t ldr(const SFD_GLOBAL char *b, unsigned ob) {
    return ld_at((const SFD_GLOBAL unsigned short *)b, ob);
  }
  static __device__ __forceinline__ float cvt(raw_t h, float s) {
    return rounded(__half2float(__ushort_as_half(h)) * s);
  }
  static __device__ __forceinline__ float ldo(const SFD_GLOBAL char *b, unsigned ob, float s) {
    return cvt(ldr(b, ob), s);
  }
  static __device__ __forceinline__ void sto(SFD_GLOBAL char *b, unsigned ob, float is, float v) {
    st_at((SFD_GLOBAL unsigned short *)b, ob, __half_as_ushort(__float2half_rn(rounded(v * is))));
  }
};

// The column (i, j) of the tile's lane and the tile's z chunk bz.  Tiles are
// 64-wide x segments of blockDim.y rows, or, in flat tile maps, runs of
// 64 * blockDim.y consecutive columns of the x-y plane in x-fastest order
// (grids whose nx + 1 is far from a multiple of 64 then leave no lane idle;
// every neighbour of the face passes comes from memory, so a run may wrap
// from one row to the next).  j > ny marks a lane past the plane.
__device__ __forceinline__ void tile_column(const TileMap &tm, int blk, int nx, int &i, int &j,
                                            int &bz) {
  int bx, by;
  decode_tile(tm, blk, bx, by, bz);
  if (tm.flat) {
    const int f = bx * kWave * (int)blockDim.y + (int)threadIdx.y * kWave + (int)threadIdx.x;
    i = f % (nx + 1);
    j = f / (nx + 1);
  } else {
    i = bx * kWave + threadIdx.x;
    j = by * blockDim.y + threadIdx.y;
  }
}

// one tile (64 x blockDim.y lanes, one z chunk) of the E face pass; T the
// field storage (fp32, or binary16 with the arrays' scales).  A lane
// addresses each z plane by a base the same for every lane (scalar
// registers) plus its byte offset in the plane, fixed over the march; the psi
// slabs likewise.  Each plane issues every load it needs (fields, media, the
// psi values of the CPML slabs) before any arithmetic, so a plane costs one
// memory round trip; the per-plane spacing and z profile are scalar loads,
// the x and y profiles fixed for the lane.
template <bool YEE, bool CPML, bool TBL, class T = float, class G = DevGrid>
__device__ __forceinline__ void epass_tile(const G &g, const TileMap &tm,
                                           const MedTab &tab, int blk, int m0 = -1) {
  typedef Store<T> S;
  typedef const SFD_GLOBAL float *cgf;
  constexpr unsigned ES = sizeof(T);
  int i, j, bz;
  tile_column(tm, blk, g.nx, i, j, bz);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (i > nx || j > ny) return;
  const int k0 = bz * tm.kc;
  const int k1 = min(k0 + tm.kc, nz + 1);
  const long long fpb = g.pxy * ES, mpb = g.pxy * 2;   // plane strides, bytes
  const unsigned lo = (unsigned)((j + 1) * (int)g.px + g.xoff + i);
  const unsigned of = lo * ES, om = lo * 2u, fy = (unsigned)g.px * ES;
  char *const bE[3] = {(char *)g.f[0], (char *)g.f[1], (char *)g.f[2]};
  char *const bH[3] = {(char *)g.f[3], (char *)g.f[4], (char *)g.f[5]};
  char *const bM[3] = {(char *)g.mi[0], (char *)g.mi[1], (char *)g.mi[2]};
  constexpr bool NEEDH = YEE || CPML;
  float hx_km = 0.0f, hy_km = 0.0f;
  if (NEEDH) {   // plane k0 - 1 of Hx, Hy
    const long long kb = (long long)k0 * fpb;
    hx_km = S::ldo(plane_base(bH[0], kb), of, g.fs[3]);
    hy_km = S::ldo(plane_base(bH[1], kb), of, g.fs[4]);
  }
  const float idyh = ((cgf)g.id[4])[j + 1];
  const float idxh = ((cgf)g.id[3])[i + 1];
  const bool cx = i < nx, cy = j < ny;   // Ex needs i<nx, Ey j<ny, Ez k<nz
  // CPML slab membership and profiles along x and y are fixed for the lane
  int qxn = 0, qyn = 0;
  bool sxn = false, syn = false;
  float bx = 0.0f, ccx = 0.0f, by = 0.0f, ccy = 0.0f;
  if (CPML) {
    sxn = slab_node(i, nx, g.L[0][0], g.L[0][1], qxn);
    syn = slab_node(j, ny, g.L[1][0], g.L[1][1], qyn);
    if (sxn) { bx = ((cgf)g.pbe[0])[i + 1]; ccx = ((cgf)g.pce[0])[i + 1]; }
    if (syn) { by = ((cgf)g.pbe[1])[j + 1]; ccy = ((cgf)g.pce[1])[j + 1]; }
  }
  const int Lx = g.L[0][0] + g.L[0][1], Ly = g.L[1][0] + g.L[1][1];
  // psi slab layouts (borderscpml.F90): byte offsets of the lane in a slab
  // plane; the plane (k, or the z slab index) adds a uniform base
  const unsigned oA = (unsigned)(qyn * nx + i) * 4u;         // psi_Exy
  const unsigned oB = (unsigned)(j * nx + i) * 4u;           // psi_Exz
  const unsigned oC = (unsigned)(j * (nx + 1) + i) * 4u;     // psi_Eyz
  const unsigned oD = (unsigned)(j * Lx + qxn) * 4u;         // psi_Eyx, psi_Ezx
  const unsigned oE = (unsigned)(qyn * (nx + 1) + i) * 4u;   // psi_Ezy
  float *const pExy = (float *)g.psi[PSI_EXY], *const pExz = (float *)g.psi[PSI_EXZ];
  float *const pEyz = (float *)g.psi[PSI_EYZ], *const pEyx = (float *)g.psi[PSI_EYX];
  float *const pEzy = (float *)g.psi[PSI_EZY], *const pEzx = (float *)g.psi[PSI_EZX];
  for (int k = k0; k < k1; ++k) {
    const bool cz = k < nz;
    const long long kb = (long long)(k + 1) * fpb, kmb = (long long)(k + 1) * mpb;
    // ---- issue the independent loads of this plane
    const float idzh = ro_uniform(g.id[5], k + 1);
    // raw loads first (binary16 storage converts after every load of the
    // plane has issued, outside the branches)
    typename S::raw_t rhx = 0, rhy = 0, rhz = 0, rhz_jm = 0, rhx_jm = 0, rhz_im = 0, rhy_im = 0;
    if (NEEDH) {
      const SFD_GLOBAL char *hxk = plane_base(bH[0], kb), *hyk = plane_base(bH[1], kb),
                            *hzk = plane_base(bH[2], kb);
      rhx = S::ldr(hxk, of); rhy = S::ldr(hyk, of); rhz = S::ldr(hzk, of);
      rhz_jm = S::ldr(hzk, of - fy); rhx_jm = S::ldr(hxk, of - fy);
      rhz_im = S::ldr(hzk, of - ES); rhy_im = S::ldr(hyk, of - ES);
    }
    SFD_GLOBAL char *const exk = plane_base(bE[0], kb), *const eyk = plane_base(bE[1], kb),
                         *const ezk = plane_base(bE[2], kb);
    // binary16: every lane loads all three components (i <= nx, j <= ny,
    // k <= nz lie inside each array's Alloc box), so that no conversion
    // waits for its load inside a branch; fp32 loads what the sweeps use
    constexpr bool ALL = sizeof(T) == 2;
    int mx = 0, my = 0, mz = 0;
    typename S::raw_t rex = 0, rey = 0, rez = 0;
    // m0 >= 0: every cell of the tile holds that medium (no media loads)
    if (ALL || cx) { mx = m0 >= 0 ? m0 : ld_at((const SFD_GLOBAL int16_t *)plane_base(bM[0], kmb), om); rex = S::ldr(exk, of); }
    if (ALL || cy) { my = m0 >= 0 ? m0 : ld_at((const SFD_GLOBAL int16_t *)plane_base(bM[1], kmb), om); rey = S::ldr(eyk, of); }
    if (ALL || cz) { mz = m0 >= 0 ? m0 : ld_at((const SFD_GLOBAL int16_t *)plane_base(bM[2], kmb), om); rez = S::ldr(ezk, of); }
    int qzn = 0;
    const bool szn = CPML && slab_node(k, nz, g.L[2][0], g.L[2][1], qzn);
    float bzp = 0.0f, ccz = 0.0f, vxy = 0.0f, vxz = 0.0f, vyz = 0.0f, vyx = 0.0f, vzy = 0.0f,
          vzx = 0.0f;
    SFD_GLOBAL float *sExy = nullptr, *sExz = nullptr, *sEyz = nullptr, *sEyx = nullptr,
                     *sEzy = nullptr, *sEzx = nullptr;
    if (CPML) {
      if (szn) { bzp = ro_uniform(g.pbe[2], k + 1); ccz = ro_uniform(g.pce[2], k + 1); }
      // slab plane bases: psi_Exy (k*Ly + q)*nx + i, psi_Exz (q*(ny+1) + j)*nx + i,
      // psi_Eyz (q*ny + j)*(nx+1) + i, psi_Eyx (k*ny + j)*Lx + q,
      // psi_Ezy (k*Ly + q)*(nx+1) + i, psi_Ezx (k*(ny+1) + j)*Lx + q
      if (Ly) {   // uniform: the y slabs exist
        sExy = plane_base(pExy, (long long)k * Ly * nx);
        sEzy = plane_base(pEzy, (long long)k * Ly * (nx + 1));
      }
      if (szn) {
        sExz = plane_base(pExz, (long long)qzn * (ny + 1) * nx);
        sEyz = plane_base(pEyz, (long long)qzn * ny * (nx + 1));
      }
      if (Lx) {
        sEyx = plane_base(pEyx, (long long)k * ny * Lx);
        sEzx = plane_base(pEzx, (long long)k * (ny + 1) * Lx);
      }
      if (cx && syn) vxy = ld_at(sExy, oA);
      if (cx && szn) vxz = ld_at(sExz, oB);
      if (cy && szn) vyz = ld_at(sEyz, oC);
      if (cy && sxn) vyx = ld_at(sEyx, oD);
      if (cz && syn) vzy = ld_at(sEzy, oE);
      if (cz && sxn) vzx = ld_at(sEzx, oD);
    }
    float hx0 = 0.0f, hy0 = 0.0f, hz0 = 0.0f, hz_jm = 0.0f, hx_jm = 0.0f, hz_im = 0.0f,
          hy_im = 0.0f;
    if (NEEDH) {
      hx0 = S::cvt(rhx, g.fs[3]); hy0 = S::cvt(rhy, g.fs[4]); hz0 = S::cvt(rhz, g.fs[5]);
      hz_jm = S::cvt(rhz_jm, g.fs[5]); hx_jm = S::cvt(rhx_jm, g.fs[3]);
      hz_im = S::cvt(rhz_im, g.fs[5]); hy_im = S::cvt(rhy_im, g.fs[4]);
    }
    const float ex = S::cvt(rex, g.fs[0]), ey = S::cvt(rey, g.fs[1]), ez = S::cvt(rez, g.fs[2]);
    // ---- compute; psi <- b*psi + c*curl, then E += +-G2*psi
    if (cx) {
      float e = ex;
      const float2 c = coef<TBL>(tab, g.g1, g.g2, mx);
      if (YEE) e = yee_upd(c.x, e, c.y, curl2(hz0 - hz_jm, idyh, hy0 - hy_km, idzh));
      if (CPML) {
        if (syn) {  // psi_Exy, left/right (borderscpml.F90:609-623)
          const float v = madd((hz0 - hz_jm) * ccy, by, vxy);
          st_at(sExy, oA, v);
          e = madd(e, c.y, v);
        }
        if (szn) {  // psi_Exz, down/up (:699-712)
          const float v = madd((hy0 - hy_km) * ccz, bzp, vxz);
          st_at(sExz, oB, v);
          e = madd(e, -c.y, v);
        }
      }
      S::sto(exk, of, g.fi[0], e);
    }
    if (cy) {
      float e = ey;
      const float2 c = coef<TBL>(tab, g.g1, g.g2, my);
      if (YEE) e = yee_upd(c.x, e, c.y, curl2(hx0 - hx_km, idzh, hz0 - hz_im, idxh));
      if (CPML) {
        if (szn) {  // psi_Eyz, down/up (:680-693)
          const float v = madd((hx0 - hx_km) * ccz, bzp, vyz);
          st_at(sEyz, oC, v);
          e = madd(e, c.y, v);
        }
        if (sxn) {  // psi_Eyx, back/front (:790-802)
          const float v = madd((hz0 - hz_im) * ccx, bx, vyx);
          st_at(sEyx, oD, v);
          e = madd(e, -c.y, v);
        }
      }
      S::sto(eyk, of, g.fi[1], e);
    }
    if (cz) {
      float e = ez;
      const float2 c = coef<TBL>(tab, g.g1, g.g2, mz);
      if (YEE) e = yee_upd(c.x, e, c.y, curl2(hy0 - hy_im, idxh, hx0 - hx_jm, idyh));
      if (CPML) {
        if (syn) {  // psi_Ezy, left/right (:630-642)
          const float v = madd((hx0 - hx_jm) * ccy, by, vzy);
          st_at(sEzy, oE, v);
          e = madd(e, -c.y, v);
        }
        if (sxn) {  // psi_Ezx, back/front (:771-783)
          const float v = madd((hy0 - hy_im) * ccx, bx, vzx);
          st_at(sEzx, oD, v);
          e = madd(e, c.y, v);
        }
      }
      S::sto(ezk, of, g.fi[2], e);
    }
    if (NEEDH) { hx_km = hx0; hy_km = hy0; }
  }
}

// one tile of the H face pass (addressing and loads as in the E pass)
template <bool YEE, bool CPML, bool TBL, class T = float, class G = DevGrid>
__device__ __forceinline__ void hpass_tile(const G &g, const TileMap &tm,
                                           const MedTab &tab, int blk, int m0 = -1) {
  typedef Store<T> S;
  typedef const SFD_GLOBAL float *cgf;
  constexpr unsigned ES = sizeof(T);
  int i, j, bz;
  tile_column(tm, blk, g.nx, i, j, bz);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (i > nx || j > ny) return;
  const int k0 = bz * tm.kc;
  const int k1 = min(k0 + tm.kc, nz + 1);
  const long long fpb = g.pxy * ES, mpb = g.pxy * 2;
  const unsigned lo = (unsigned)((j + 1) * (int)g.px + g.xoff + i);
  const unsigned of = lo * ES, om = lo * 2u, fy = (unsigned)g.px * ES;
  char *const bE[3] = {(char *)g.f[0], (char *)g.f[1], (char *)g.f[2]};
  char *const bH[3] = {(char *)g.f[3], (char *)g.f[4], (char *)g.f[5]};
  char *const bM[3] = {(char *)g.mi[3], (char *)g.mi[4], (char *)g.mi[5]};
  const float idye = ((cgf)g.id[1])[j + 1];
  const float idxe = ((cgf)g.id[0])[i + 1];
  float ex_c, ey_c;
  {  // plane k0 of Ex, Ey
    const long long kb = (long long)(k0 + 1) * fpb;
    ex_c = S::ldo(plane_base(bE[0], kb), of, g.fs[0]);
    ey_c = S::ldo(plane_base(bE[1], kb), of, g.fs[1]);
  }
  int qxc = 0, qyc = 0;
  bool sxc = false, syc = false;
  float bx = 0.0f, ccx = 0.0f, by = 0.0f, ccy = 0.0f;
  if (CPML) {
    // slab_cell also answers true at c = n on an axis without layers (whose
    // profiles are then absent): the x cells count for Hy and Hz only (i < nx),
    // the y cells for Hx and Hz (j < ny), the z cells for Hx and Hy (k < nz)
    sxc = i < nx && slab_cell(i, nx, g.L[0][0], g.L[0][1], qxc);
    syc = j < ny && slab_cell(j, ny, g.L[1][0], g.L[1][1], qyc);
    if (sxc) { bx = ((cgf)g.pbm[0])[i + 1]; ccx = ((cgf)g.pcm[0])[i + 1]; }
    if (syc) { by = ((cgf)g.pbm[1])[j + 1]; ccy = ((cgf)g.pcm[1])[j + 1]; }
  }
  const int Lx = g.L[0][0] + g.L[0][1], Ly = g.L[1][0] + g.L[1][1];
  const unsigned oA = (unsigned)(qyc * (nx + 1) + i) * 4u;   // psi_Hxy
  const unsigned oB = (unsigned)(j * (nx + 1) + i) * 4u;     // psi_Hxz
  const unsigned oC = (unsigned)(j * nx + i) * 4u;           // psi_Hyz
  const unsigned oD = (unsigned)(j * Lx + qxc) * 4u;         // psi_Hyx, psi_Hzx
  const unsigned oE = (unsigned)(qyc * nx + i) * 4u;         // psi_Hzy
  float *const pHxy = (float *)g.psi[PSI_HXY], *const pHxz = (float *)g.psi[PSI_HXZ];
  float *const pHyz = (float *)g.psi[PSI_HYZ], *const pHyx = (float *)g.psi[PSI_HYX];
  float *const pHzy = (float *)g.psi[PSI_HZY], *const pHzx = (float *)g.psi[PSI_HZX];
  for (int k = k0; k < k1; ++k) {
    const bool ax = j < ny && k < nz;   // Hx sweep: x[0,nx] y[0,ny-1] z[0,nz-1]
    const bool ay = i < nx && k < nz;   // Hy sweep: x[0,nx-1] y[0,ny] z[0,nz-1]
    const bool az = i < nx && j < ny;   // Hz sweep: x[0,nx-1] y[0,ny-1] z[0,nz]
    const long long kb = (long long)(k + 1) * fpb, kmb = (long long)(k + 1) * mpb;
    const float idze = ro_uniform(g.id[2], k + 1);
    const SFD_GLOBAL char *exn = plane_base(bE[0], kb + fpb), *eyn = plane_base(bE[1], kb + fpb);
    const SFD_GLOBAL char *exk = plane_base(bE[0], kb), *eyk = plane_base(bE[1], kb),
                          *ezk = plane_base(bE[2], kb);
    // raw loads first (converted once every load of the plane has issued)
    const typename S::raw_t rex_kp = S::ldr(exn, of), rey_kp = S::ldr(eyn, of);
    const typename S::raw_t rez_c = S::ldr(ezk, of);
    const typename S::raw_t rez_jp = S::ldr(ezk, of + fy), rex_jp = S::ldr(exk, of + fy);
    const typename S::raw_t rez_ip = S::ldr(ezk, of + ES), rey_ip = S::ldr(eyk, of + ES);
    SFD_GLOBAL char *const hxk = plane_base(bH[0], kb), *const hyk = plane_base(bH[1], kb),
                         *const hzk = plane_base(bH[2], kb);
    // binary16: all three components for every lane (inside each Alloc box)
    constexpr bool ALL = sizeof(T) == 2;
    int mx = 0, my = 0, mz = 0;
    typename S::raw_t rhx = 0, rhy = 0, rhz = 0;
    if (ALL || ax) { mx = m0 >= 0 ? m0 : ld_at((const SFD_GLOBAL int16_t *)plane_base(bM[0], kmb), om); rhx = S::ldr(hxk, of); }
    if (ALL || ay) { my = m0 >= 0 ? m0 : ld_at((const SFD_GLOBAL int16_t *)plane_base(bM[1], kmb), om); rhy = S::ldr(hyk, of); }
    if (ALL || az) { mz = m0 >= 0 ? m0 : ld_at((const SFD_GLOBAL int16_t *)plane_base(bM[2], kmb), om); rhz = S::ldr(hzk, of); }
    int qzc = 0;
    const bool szc = CPML && k < nz && slab_cell(k, nz, g.L[2][0], g.L[2][1], qzc);
    float bzp = 0.0f, ccz = 0.0f, vxy = 0.0f, vxz = 0.0f, vyz = 0.0f, vyx = 0.0f, vzy = 0.0f,
          vzx = 0.0f;
    SFD_GLOBAL float *sHxy = nullptr, *sHxz = nullptr, *sHyz = nullptr, *sHyx = nullptr,
                     *sHzy = nullptr, *sHzx = nullptr;
    if (CPML) {
      if (szc) { bzp = ro_uniform(g.pbm[2], k + 1); ccz = ro_uniform(g.pcm[2], k + 1); }
      // slab plane bases: psi_Hxy (k*Ly + q)*(nx+1) + i, psi_Hxz (q*ny + j)*(nx+1) + i,
      // psi_Hyz (q*(ny+1) + j)*nx + i, psi_Hyx (k*(ny+1) + j)*Lx + q,
      // psi_Hzy (k*Ly + q)*nx + i, psi_Hzx (k*ny + j)*Lx + q
      if (Ly) {   // uniform: the y slabs exist
        sHxy = plane_base(pHxy, (long long)k * Ly * (nx + 1));
        sHzy = plane_base(pHzy, (long long)k * Ly * nx);
      }
      if (szc) {
        sHxz = plane_base(pHxz, (long long)qzc * ny * (nx + 1));
        sHyz = plane_base(pHyz, (long long)qzc * (ny + 1) * nx);
      }
      if (Lx) {
        sHyx = plane_base(pHyx, (long long)k * (ny + 1) * Lx);
        sHzx = plane_base(pHzx, (long long)k * ny * Lx);
      }
      if (ax && syc) vxy = ld_at(sHxy, oA);
      if (ax && szc) vxz = ld_at(sHxz, oB);
      if (ay && szc) vyz = ld_at(sHyz, oC);
      if (ay && sxc) vyx = ld_at(sHyx, oD);
      if (az && syc) vzy = ld_at(sHzy, oE);
      if (az && sxc) vzx = ld_at(sHzx, oD);
    }
    const float ex_kp = S::cvt(rex_kp, g.fs[0]), ey_kp = S::cvt(rey_kp, g.fs[1]);
    const float ez_c = S::cvt(rez_c, g.fs[2]);
    const float ez_jp = S::cvt(rez_jp, g.fs[2]), ex_jp = S::cvt(rex_jp, g.fs[0]);
    const float ez_ip = S::cvt(rez_ip, g.fs[2]), ey_ip = S::cvt(rey_ip, g.fs[1]);
    const float hx = S::cvt(rhx, g.fs[3]), hy = S::cvt(rhy, g.fs[4]), hz = S::cvt(rhz, g.fs[5]);
    if (ax) {
      const float2 c = coef<TBL>(tab, g.gm1, g.gm2, mx);
      float h = hx;
      if (YEE) h = yee_upd(c.x, h, c.y, curl2(ey_kp - ey_c, idze, ez_jp - ez_c, idye));
      if (CPML) {
        if (syc) {  // psi_Hxy (borderscpml.F90:911-921)
          const float v = madd((ez_jp - ez_c) * ccy, by, vxy);
          st_at(sHxy, oA, v);
          h = madd(h, -c.y, v);
        }
        if (szc) {  // psi_Hxz (:1021-1031)
          const float v = madd((ey_kp - ey_c) * ccz, bzp, vxz);
          st_at(sHxz, oB, v);
          h = madd(h, c.y, v);
        }
      }
      S::sto(hxk, of, g.fi[3], h);
    }
    if (ay) {
      const float2 c = coef<TBL>(tab, g.gm1, g.gm2, my);
      float h = hy;
      if (YEE) h = yee_upd(c.x, h, c.y, curl2(ez_ip - ez_c, idxe, ex_kp - ex_c, idze));
      if (CPML) {
        if (szc) {  // psi_Hyz (:1001-1011)
          const float v = madd((ex_kp - ex_c) * ccz, bzp, vyz);
          st_at(sHyz, oC, v);
          h = madd(h, -c.y, v);
        }
        if (sxc) {  // psi_Hyx (:1111-1121)
          const float v = madd((ez_ip - ez_c) * ccx, bx, vyx);
          st_at(sHyx, oD, v);
          h = madd(h, c.y, v);
        }
      }
      S::sto(hyk, of, g.fi[4], h);
    }
    if (az) {
      const float2 c = coef<TBL>(tab, g.gm1, g.gm2, mz);
      float h = hz;
      if (YEE) h = yee_upd(c.x, h, c.y, curl2(ex_jp - ex_c, idye, ey_ip - ey_c, idxe));
      if (CPML) {
        if (syc) {  // psi_Hzy (:931-941)
          const float v = madd((ex_jp - ex_c) * ccy, by, vzy);
          st_at(sHzy, oE, v);
          h = madd(h, c.y, v);
        }
        if (sxc) {  // psi_Hzx (:1091-1101)
          const float v = madd((ey_ip - ey_c) * ccx, bx, vzx);
          st_at(sHzx, oD, v);
          h = madd(h, -c.y, v);
        }
      }
      S::sto(hzk, of, g.fi[5], h);
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
  long long p = (long long)(k0 + 1) * pxy + (long long)(j + 1) * px + g.xoff + i0;
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
  long long p = (long long)(k0 + 1) * pxy + (long long)(j + 1) * px + g.xoff + i0;
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

template <bool YEE, bool CPML, bool TBL, class T = float>
__global__ void __launch_bounds__(256) k_epass(DevGrid g, TileMap tm, const int *tmed) {
  extern __shared__ float2 lds_tab[];
  const MedTab tab = stage_tables<TBL, false>(g, g.g1, g.g2, lds_tab);
  const int m0 = tmed ? __builtin_amdgcn_readfirstlane(tmed[blockIdx.x]) : -1;
  epass_tile<YEE, CPML, TBL, T>(g, tm, tab, blockIdx.x, m0);
}

template <bool YEE, bool CPML, bool TBL, class T = float>
__global__ void __launch_bounds__(256) k_hpass(DevGrid g, TileMap tm, const int *tmed) {
  extern __shared__ float2 lds_tab[];
  const MedTab tab = stage_tables<TBL, false>(g, g.gm1, g.gm2, lds_tab);
  const int m0 = tmed ? __builtin_amdgcn_readfirstlane(tmed[blockIdx.x]) : -1;
  hpass_tile<YEE, CPML, TBL, T>(g, tm, tab, blockIdx.x, m0);
}

// One block per tile of a one-cell pass map: the medium index that all six
// media arrays hold on every cell the tile's lanes own (columns i <= nx,
// j <= ny, the planes of its z chunk), or -1 where they differ.  The tiles
// are decoded as the passes decode them.
__global__ void __launch_bounds__(256) k_pass_tile_media(DevGrid g, TileMap tm, int *out) {
  int bx, by, bz;
  decode_tile(tm, blockIdx.x, bx, by, bz);
  const int nx1 = g.nx + 1, ny1 = g.ny + 1;
  const int k0 = bz * tm.kc, k1 = min(k0 + tm.kc, g.nz + 1);
  long long c0 = 0, nc;
  int i0 = 0, j0 = 0, wi = 1;
  if (tm.flat) {
    c0 = (long long)bx * kWave * tm.bdy;
    const long long ce = min(c0 + (long long)kWave * tm.bdy, (long long)nx1 * ny1);
    nc = ce - c0;
  } else {
    i0 = bx * kWave;
    j0 = by * tm.bdy;
    wi = min(kWave, nx1 - i0);
    const int wj = min(tm.bdy, ny1 - j0);
    nc = wi > 0 && wj > 0 ? (long long)wi * wj : 0;
  }
  if (nc <= 0 || k1 <= k0) {
    if (threadIdx.x == 0) out[blockIdx.x] = -1;
    return;
  }
  __shared__ int bad;
  auto at = [&](int i, int j, int k) {
    return (long long)(k + 1) * g.pxy + (long long)(j + 1) * g.px + g.xoff + i;
  };
  int ia = i0, ja = j0;
  if (tm.flat) { ia = (int)(c0 % nx1); ja = (int)(c0 / nx1); }
  const int m0 = g.mi[0][at(ia, ja, k0)];
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  const long long n = nc * (k1 - k0);
  int b = 0;
  for (long long q = threadIdx.x; q < n && !b; q += blockDim.x) {
    const long long col = q % nc;
    const int k = k0 + (int)(q / nc);
    int i, j;
    if (tm.flat) { i = (int)((c0 + col) % nx1); j = (int)((c0 + col) / nx1); }
    else { i = i0 + (int)(col % wi); j = j0 + (int)(col / wi); }
    const long long p = at(i, j, k);
#pragma unroll
    for (int c = 0; c < 6; ++c) b |= g.mi[c][p] != m0;
  }
  if (b) atomicOr(&bad, 1);
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = bad ? -1 : m0;
}

int launch_pass_tile_media(const DevGrid &g, const TileMap &tm, int *out, void *stream) {
  if (tm.nblk <= 0) return 0;
  hipLaunchKernelGGL(k_pass_tile_media, dim3((unsigned)tm.nblk), dim3(256), 0,
                     (hipStream_t)stream, g, tm, out);
  return (int)hipGetLastError();
}

// ---- batches of independent grids ----------------------------------------
// One persistent launch per half step covers every tile of every grid of the
// batch: block b takes tiles b, b + gridDim.x, ...; the grid of a tile is
// found by binary search in the prefix of tile counts, its media tables are
// staged in LDS before the tile runs.
// a tile map read word by word through the constant address space (scalar
// loads; the maps never change during a launch)
__device__ __forceinline__ TileMap const_tilemap(const TileMap *t) {
  static_assert(sizeof(TileMap) % sizeof(int) == 0, "TileMap holds ints only");
  typedef const __attribute__((address_space(4))) int CInt;
  TileMap m;
  int *o = reinterpret_cast<int *>(&m);
  for (int q = 0; q < (int)(sizeof(TileMap) / sizeof(int)); ++q) o[q] = ((CInt *)t)[q];
  return m;
}

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
    const int b = __builtin_amdgcn_readfirstlane(batch_grid(bd.start, bd.ngrid, t));
    const DevGrid &g = bd.grids[b];
    // the one-cell passes read the grid's descriptor through the constant
    // address space: scalar loads into scalar registers (it never changes
    // during the launch), not vector loads into some 40 vector registers
    typedef __attribute__((address_space(4))) DevGrid CDevGrid;
    const CDevGrid &cg = *(const CDevGrid *)(bd.grids + b);
    const TileMap tm = const_tilemap(bd.tms + b);
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
      // a tile whose cells all hold one medium skips the media loads
      const int m0 = bd.tmed ? __builtin_amdgcn_readfirstlane(bd.tmed[t]) : -1;
      if (H) hpass_tile<true, true, true, T, CDevGrid>(cg, tm, tab, blk, m0);
      else epass_tile<true, true, true, T, CDevGrid>(cg, tm, tab, blk, m0);
    }
  }
}

int launch_pass_batch(const BatchDev &bd, int half, int blocks, int rows, int vec, size_t lds,
                      int fp16, void *stream) {
  if (bd.total <= 0) return 0;
  const dim3 blk(kWave, (unsigned)rows);
  hipStream_t s = (hipStream_t)stream;
  if (fp16 && vec == 4) {
    if (half) hipLaunchKernelGGL((k_pass_batch<true, __half, 4>), dim3(blocks), blk, lds, s, bd);
    else hipLaunchKernelGGL((k_pass_batch<false, __half, 4>), dim3(blocks), blk, lds, s, bd);
  } else if (fp16) {
    if (half) hipLaunchKernelGGL((k_pass_batch<true, __half, 1>), dim3(blocks), blk, lds, s, bd);
    else hipLaunchKernelGGL((k_pass_batch<false, __half, 1>), dim3(blocks), blk, lds, s, bd);
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

size_t ade_lds_of(const DevGrid &g, int side) { return ade_lds_bytes(g, side); }

int launch_ade_many(const DevGrid *grids, const int *list, int n, int side, size_t lds,
                    int tbl, long long nmax, void *stream) {
  if (n <= 0 || nmax <= 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  const long long want = (nmax + 255) / 256;
  const dim3 grid((unsigned)(want < 1024 ? want : 1024), 3, (unsigned)n);
  if (side == 0) {
    if (tbl) hipLaunchKernelGGL((k_ade_many<true, 0>), grid, dim3(256), lds, s, grids, list);
    else hipLaunchKernelGGL((k_ade_many<false, 0>), grid, dim3(256), 0, s, grids, list);
  } else {
    if (tbl) hipLaunchKernelGGL((k_ade_many<true, 1>), grid, dim3(256), lds, s, grids, list);
    else hipLaunchKernelGGL((k_ade_many<false, 1>), grid, dim3(256), 0, s, grids, list);
  }
  return (int)hipGetLastError();
}

int launch_e(const DevGrid &g, const TileMap &tm, const TileMap &tm4, const int *tmed, int yee,
             int cpml, int ade, void *stream) {
  const dim3 blk(kWave, (unsigned)tm.bdy);
  hipStream_t s = (hipStream_t)stream;
  const int nb = tm.nblk;
  const size_t lds = e_lds_bytes(g);
  const bool tbl = g.nmedia + 1 <= kLdsMedia;
#define SFD_E(Y, C)                                                            \
  do {                                                                         \
    if (tbl)                                                                   \
      hipLaunchKernelGGL((k_epass<Y, C, true>), dim3(nb), blk, lds, s, g, tm, tmed); \
    else                                                                       \
      hipLaunchKernelGGL((k_epass<Y, C, false>), dim3(nb), blk, 0, s, g, tm, tmed); \
  } while (0)
  if (yee && cpml && tbl && tm4.vec == 4) {
    // fused pass, four cells per lane (the CPML variant also serves grids
    // without CPML)
    if (g.fp16) hipLaunchKernelGGL(k_epass4<__half>, dim3(tm4.nblk), blk, lds, s, g, tm4);
    else hipLaunchKernelGGL(k_epass4<float>, dim3(tm4.nblk), blk, lds, s, g, tm4);
  } else if (g.fp16 && yee && cpml && tbl) {
    // binary16 storage, one cell per lane (the set-up admits binary16 only
    // with the fused Yee + CPML pass and LDS tables)
    hipLaunchKernelGGL((k_epass<true, true, true, __half>), dim3(nb), blk, lds, s, g, tm, tmed);
  } else if (yee && cpml) SFD_E(true, true);
  else if (yee) SFD_E(true, false);
  else if (cpml) SFD_E(false, true);
#undef SFD_E
  if (ade) launch_ade(g, 0, s);
  return (int)hipGetLastError();
}

int launch_h(const DevGrid &g, const TileMap &tm, const TileMap &tm4, const int *tmed, int yee,
             int cpml, int ade, void *stream) {
  const dim3 blk(kWave, (unsigned)tm.bdy);
  hipStream_t s = (hipStream_t)stream;
  const int nb = tm.nblk;
  const bool tbl = g.nmedia + 1 <= kLdsMedia;
  const size_t lds = tbl ? (size_t)(g.nmedia + 1) * sizeof(float2) : 0;
#define SFD_H(Y, C)                                                            \
  do {                                                                         \
    if (tbl)                                                                   \
      hipLaunchKernelGGL((k_hpass<Y, C, true>), dim3(nb), blk, lds, s, g, tm, tmed); \
    else                                                                       \
      hipLaunchKernelGGL((k_hpass<Y, C, false>), dim3(nb), blk, 0, s, g, tm, tmed); \
  } while (0)
  if (yee && cpml && tbl && tm4.vec == 4) {
    if (g.fp16) hipLaunchKernelGGL(k_hpass4<__half>, dim3(tm4.nblk), blk, lds, s, g, tm4);
    else hipLaunchKernelGGL(k_hpass4<float>, dim3(tm4.nblk), blk, lds, s, g, tm4);
  } else if (g.fp16 && yee && cpml && tbl) {
    hipLaunchKernelGGL((k_hpass<true, true, true, __half>), dim3(nb), blk, lds, s, g, tm, tmed);
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
// A block of 64 x R threads marches the z chunks of a tile.  Lanes 0..62 and
// rows 0..R-2 own their cells; lane 63 and row R-1 overlap the neighbouring
// tiles: their E values feed the owned H cells and are not stored, and the
// plane above the chunk is computed the same way.  E values travel to the
// x - 1 neighbour by a cross-lane shuffle and to the y - 1 neighbour through
// LDS; H values of step n are read from memory (gi is never written during
// the step).


// the descriptor pointer made opaque to the compiler once per plane: what is
// read through it is read where it is used (slab and dispersive cells), not
// hoisted out of the march into scalar registers that the plain cells pay for
typedef const __attribute__((address_space(4))) StepAux CStepAux;
__device__ __forceinline__ CStepAux *per_plane(const StepAux *p) {
  CStepAux *q = (CStepAux *)p;
  asm volatile("" : "+s"(q));
  return q;
}

// the base of output field c: from the descriptor (CPML, ADE, SRC classes) or
// the kernel argument (plain tiles)
#define fo_at(a, go, c) ((CPML || ADE || SRC) ? (a)->fo[c] : (go).f[c])

// the values one plane of the march loads for the E half of the plane
struct PlaneLoad {
  float hx, hy, hz, hz_jm, hx_jm, hz_im, hy_im, ex, ey, ez;
  int mx, my, mz, mhx, mhy, mhz;
};
// FI, MI: the field and media array bases, either held in scalar registers
// across the march (the plain tiles, which have registers to spare) or read
// per plane from the step descriptor (the other classes, whose own state
// would otherwise push them into spills)
template <class FI, class MI>
__device__ __forceinline__ void load_plane(PlaneLoad &L, const FI &fi, const MI &mi, long long P,
                                           unsigned ob, unsigned obx, unsigned oby,
                                           unsigned om) {
  const SFD_GLOBAL float *hxk = plane_base(fi[3], P), *hyk = plane_base(fi[4], P),
                         *hzk = plane_base(fi[5], P);
  L.hx = ld_at(hxk, ob); L.hy = ld_at(hyk, ob); L.hz = ld_at(hzk, ob);
  L.hz_jm = ld_at(hzk, oby); L.hx_jm = ld_at(hxk, oby);
  L.hz_im = ld_at(hzk, obx); L.hy_im = ld_at(hyk, obx);
  L.ex = ld_at(plane_base(fi[0], P), ob); L.ey = ld_at(plane_base(fi[1], P), ob);
  L.ez = ld_at(plane_base(fi[2], P), ob);
  L.mx = ld_at(plane_base(mi[0], P), om); L.my = ld_at(plane_base(mi[1], P), om);
  L.mz = ld_at(plane_base(mi[2], P), om);
  L.mhx = ld_at(plane_base(mi[3], P), om); L.mhy = ld_at(plane_base(mi[4], P), om);
  L.mhz = ld_at(plane_base(mi[5], P), om);
}

// CPML psi of one correction: the new value from the step-n value old,
// written to the other copy when wr
__device__ __forceinline__ float psi_new(float old, float *po, bool wr, float b, float d,
                                         float c) {
  const float v = madd(d * c, b, old);
  if (wr) *po = v;
  return v;
}

// The pole state of one field (E or H) at one plane k, per component: the
// lane's offset in the xy extent of the component's state box (~0u outside),
// fixed for the lane, and the plane's offset zo in the box.
__device__ __forceinline__ void pole_xy(const AdeSet &A, int i, int j, unsigned *axy) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int *b = A.abox[c];
    axy[c] = (i >= b[0] && i < b[0] + b[3] && j >= b[1] && j < b[1] + b[4])
                 ? (unsigned)((j - b[1]) * b[3] + (i - b[0]))
                 : ~0u;
  }
}
template <class AS>
__device__ __forceinline__ bool pole_plane(const AS &A, int c, int k, long long &zo) {
  const auto &b = A.abox[c];
  zo = (long long)(k - b[2]) * b[3] * b[4];
  return k >= b[2] && k < b[2] + b[5];
}
// load the previous field and the first kPre pole states of edge zo + axy
template <class AS>
__device__ __forceinline__ void pole_load(const AS &A, int c, long long zo, unsigned axy,
                                          float &ep, jstate_t *pj, bool with_prev = true) {
  if (with_prev) ep = (A.prev[c] + zo)[axy];
  const jstate_t *J = reinterpret_cast<const jstate_t *>(A.jst[c]) + zo;
#pragma unroll
  for (int q = 0; q < kPre; ++q)
    if (q < A.maxp) pj[q] = (J + q * A.nstate[c])[axy];
}
// the new field of that edge (slot: its dispersive medium); the rest of its
// poles are read here, the new state is written to ao when wr, the new
// previous field when wr and wrp
template <class AS>
__device__ __forceinline__ float pole_new(const AS &ai, const AS &ao, int c,
                                          long long zo, unsigned axy, bool wr, bool wrp,
                                          double dt, const MedTab &t, int slot, float ep,
                                          const jstate_t *pj, float e) {
  const long long N = ai.nstate[c];
  const jstate_t *Ji = reinterpret_cast<const jstate_t *>(ai.jst[c]) + zo;
  jstate_t *Jo = reinterpret_cast<jstate_t *>(ao.jst[c]) + zo;
  e = ade_core(t.acf + (size_t)slot * kPMax * 6, t.anp[slot], ep, dt, e,
               [&](int q) { return q < kPre ? pj[q] : (Ji + q * N)[axy]; },
               [&](int q, jstate_t v) {
                 if (wr) (Jo + q * N)[axy] = v;
               });
  if (wr && wrp) (ao.prev[c] + zo)[axy] = e;
  return e;
}

// nodal cells of one half acting on component comp at field offset p, in the
// reference's order
__device__ __noinline__ float step_nodal(const StepSrc *sp, int half, int comp, long long p,
                                         float v, int n) {
  const StepSrc &sd = *sp;
  for (int q = 0; q < sd.ncell[half]; ++q) {
    const NodalCell &c = sd.cells[half][q];
    if (c.comp != comp || c.p != p || c.skip) continue;
    v = nodal_apply(v, c, sd.values[(long long)n * sd.nsrc + c.src]);
  }
  return v;
}

// can a wave of row j, columns [x0, x0+63] hold a nodal cell of the half
__device__ __forceinline__ bool step_nd_xy(const StepDev &sd, int half, int x0, int j) {
  const int *b = sd.cellbox[half];
  return sd.ncell[half] > 0 && j >= b[1] && j <= b[4] && x0 + kWave - 1 >= b[0] && x0 <= b[3];
}

// the incident field of face f at cell (i, j, k), from the buffer filled
// before the step
__device__ __forceinline__ float face_inc(const float *inc, const PwFace &f, int i, int j,
                                          int k) {
  const int cc[3] = {i, j, k};
  const int ua = f.axis == 0 ? 1 : 0, va = f.axis == 2 ? 1 : 2;
  const long long nu = f.hi[ua] - f.lo[ua] + 1;
  return inc[f.inc + (long long)(cc[va] - f.lo[va]) * nu + (cc[ua] - f.lo[ua])];
}
// the faces in bits applied in order to v; the first face's incident value
// was loaded with the plane (first)
__device__ __forceinline__ float faces_apply(const PwFace *F, unsigned long long bits,
                                             float first, const float *inc, float G, int i,
                                             int j, int k, float v) {
  int q = __builtin_ctzll(bits);
  v = pw_apply(v, F[q].sign, G, first, F[q].id);
  bits &= bits - 1;
  while (bits) {
    q = __builtin_ctzll(bits);
    bits &= bits - 1;
    v = pw_apply(v, F[q].sign, G, face_inc(inc, F[q], i, j, k), F[q].id);
  }
  return v;
}

// does face f of a half cross column (i, j) of the grid (any plane)
__device__ __forceinline__ bool face_column(const PwFace &f, int i, int j) {
  if (f.axis == 0) return i == f.fixed && j >= f.lo[1] && j <= f.hi[1];
  if (f.axis == 1) return j == f.fixed && i >= f.lo[0] && i <= f.hi[0];
  return i >= f.lo[0] && i <= f.hi[0] && j >= f.lo[1] && j <= f.hi[1];
}
// does face f cross plane k
__device__ __forceinline__ bool face_plane(const PwFace &f, int k) {
  return f.axis == 2 ? k == f.fixed : (k >= f.lo[2] && k <= f.hi[2]);
}

// LDS of one block: the E and H media tables, the E exchange rows, and for
// grids with plane waves the faces of both halves, their per-plane masks over
// the chunk's kc + 1 planes and their per-component masks
static __host__ __device__ __forceinline__ size_t step_lds_faces(size_t base, int npw, int kc,
                                                                size_t *offz, size_t *offc) {
  const size_t f = base + (size_t)npw * sizeof(PwFace);
  const size_t z = (f + 7) & ~(size_t)7;
  if (offz) *offz = z;
  const size_t c = z + (size_t)2 * (kc + 1) * 8;
  if (offc) *offc = c;
  return c + 6 * 8;
}

static __host__ __device__ __forceinline__ size_t step_tab_bytes(int nmedia, int nd) {
  return (((size_t)(nmedia + 1) * 12 + (size_t)nd * 4 + 7) & ~(size_t)7) +
         (size_t)nd * kPMax * 6 * sizeof(coef_t);
}

// Tiles are launched in classes: CPML, tiles that touch a CPML slab; ADE,
// tiles that meet the pole-state box of some component; SRC, tiles that
// touch a TF/SF face or a nodal cell.  A class's kernel carries only the code
// its tiles need, so interior tiles run the plain Yee march.
// XW: lanes per grid row.  64: a wave holds one row of 64 columns; 32: a wave
// holds two rows of 32 columns (half-width tiles, for pieces of at most 31
// columns such as the x CPML slabs, which would leave most of a 64-lane row
// idle); the block's R waves then hold 2R rows.
template <int R, bool CPML, bool ADE, bool SRC, class MT, int XW>
__device__ __forceinline__ void step_tile(const StepDev &sd, const StepTile tl) {
  static_assert(XW == 64 || XW == 32, "a row is a whole or a half wave");
  constexpr int RR = R * (kWave / XW);   // grid rows per block
  extern __shared__ float2 lds[];
  const DevGrid &gi = sd.gi;
  const DevGrid &go = sd.go;
  const size_t offh = step_tab_bytes(gi.nmedia, gi.ade[0].ndisp);
  const size_t offx = offh + step_tab_bytes(gi.nmedia, gi.ade[1].ndisp);
  const MedTab te = stage_tables<true, ADE>(gi, gi.g1, gi.g2, lds, &gi.ade[0]);
  const MedTab th = stage_tables<true, ADE>(
      gi, gi.gm1, gi.gm2, reinterpret_cast<float2 *>(reinterpret_cast<char *>(lds) + offh),
      &gi.ade[1]);
  float2 *xch = reinterpret_cast<float2 *>(reinterpret_cast<char *>(lds) + offx);
  const int x0 = tl.x & 0xfffff, wx = tl.x >> 20;
  const int y0 = tl.y & 0xfffff, wy = tl.y >> 20;
  const int k0 = tl.k0;
  // TF/SF faces of both halves in LDS (E faces, then H faces), with the mask
  // of the faces crossing each plane of the chunk (E: k0..k0+kc, H: the same
  // numbering) and the mask of each component's faces
  PwFace *lf = nullptr;
  unsigned long long *lzm = nullptr, *lcm = nullptr;
  const int np0 = SRC ? sd.npw[0] : 0, np1 = SRC ? sd.npw[1] : 0;
  const int kc1 = sd.tm.kc + 1;
  if (SRC && np0 + np1 > 0) {
    size_t offz = 0, offc = 0;
    step_lds_faces(offx + (size_t)2 * R * kWave * sizeof(float2), np0 + np1, sd.tm.kc, &offz,
                   &offc);
    lf = reinterpret_cast<PwFace *>(reinterpret_cast<char *>(lds) + offx +
                                    (size_t)2 * R * kWave * sizeof(float2));
    lzm = reinterpret_cast<unsigned long long *>(reinterpret_cast<char *>(lds) + offz);
    lcm = reinterpret_cast<unsigned long long *>(reinterpret_cast<char *>(lds) + offc);
    const StepSrc &S = *sd.src;
    const int tid = threadIdx.y * kWave + threadIdx.x, nt = kWave * R;
    for (int q = tid; q < np0 + np1; q += nt) lf[q] = q < np0 ? S.pwf[0][q] : S.pwf[1][q - np0];
    for (int u = tid; u < 2 * kc1; u += nt) {
      const int hf = u / kc1, k = k0 + u % kc1;
      const PwFace *F = S.pwf[hf];
      unsigned long long m = 0;
      for (int q = 0; q < (hf ? np1 : np0); ++q)
        if (face_plane(F[q], k)) m |= 1ull << q;
      lzm[u] = m;
    }
    if (tid < 6) {
      const int hf = tid / 3, comp = tid;
      const PwFace *F = S.pwf[hf];
      unsigned long long m = 0;
      for (int q = 0; q < (hf ? np1 : np0); ++q)
        if (F[q].comp == comp) m |= 1ull << q;
      lcm[tid] = m;
    }
  }
  __syncthreads();

  // lane: the column in the tile's row; row: the tile's row (a wave holds
  // kWave / XW consecutive rows)
  const int lane = (int)threadIdx.x % XW, row = (int)threadIdx.y * (kWave / XW) + (int)threadIdx.x / XW;
  // lanes and rows past the halo work on the halo column and row again: their
  // results are never used, and their loads hit the halo's lines instead of
  // fetching columns and rows the tile does not own
  const int i = x0 + min(lane, wx), j = y0 + min(row, wy);
  const int nx = gi.nx, ny = gi.ny, nz = gi.nz;
  const bool in = i <= nx && j <= ny;
  // the tile owns lanes [0, wx) (wx <= XW - 1) and rows [0, wy) (wy <= RR -
  // 1); lane wx and row wy are its halo
  const bool own = in && lane < wx && row < wy;
  // lanes past the grid's edge address the edge cell (their results are
  // never used): every load and table index stays inside its array
  const int ic = min(i, nx), jc = min(j, ny);
  const int k1 = min((int)((unsigned)tl.k1 & kTileK1Mask), nz + 1);
  const long long pxy = gi.pxy;
  const unsigned pxb = (unsigned)gi.px * 4u;
  // byte offsets inside a plane: the cell and its x - 1 and y - 1 neighbours;
  // om: the cell in the int16 media arrays
  const unsigned ob = (unsigned)(jc + 1) * pxb + (unsigned)(gi.xoff + ic) * 4u;
  const unsigned obx = ob - 4u, oby = ob - pxb, om = ob >> 1;
  const bool ex_on = in && i < nx, ey_on = in && j < ny;
  const bool hx_on = own && j < ny, hy_on = own && i < nx, hz_on = own && i < nx && j < ny;
  const StepAux *ax = sd.aux;
  // CPML: slab coordinates along x and y and the x/y profile coefficients,
  // fixed for the lane (node positions for E, cell positions for H)
  int qxn = 0, qyn = 0, qxc = 0, qyc = 0;
  bool sxn = false, syn = false, sxc = false, syc = false;
  float bey = 0.f, cey = 0.f, bex = 0.f, cex = 0.f, bmy = 0.f, cmy = 0.f, bmx = 0.f, cmx = 0.f;
  long long Lx = 0, Ly = 0;
  const int Lz0 = gi.L[2][0], Lz1 = gi.L[2][1];
  if (CPML) {
    sxn = slab_node(i, nx, gi.L[0][0], gi.L[0][1], qxn);
    syn = slab_node(j, ny, gi.L[1][0], gi.L[1][1], qyn);
    sxc = slab_cell(i, nx, gi.L[0][0], gi.L[0][1], qxc);
    syc = slab_cell(j, ny, gi.L[1][0], gi.L[1][1], qyc);
    Lx = gi.L[0][0] + gi.L[0][1];
    Ly = gi.L[1][0] + gi.L[1][1];
    if (syn) { bey = gi.pbe[1][jc + 1]; cey = gi.pce[1][jc + 1]; }
    if (sxn) { bex = gi.pbe[0][ic + 1]; cex = gi.pce[0][ic + 1]; }
    if (syc) { bmy = gi.pbm[1][jc + 1]; cmy = gi.pcm[1][jc + 1]; }
    if (sxc) { bmx = gi.pbm[0][ic + 1]; cmx = gi.pcm[0][ic + 1]; }
  }
  const float idyh = gi.id[4][jc + 1], idxh = gi.id[3][ic + 1];
  const float idye = gi.id[1][jc + 1], idxe = gi.id[0][ic + 1];
  const float *idzh_t = gi.id[5], *idze_t = gi.id[2];
  const int n = SRC ? *sd.dn : 0;
  // the faces of each half crossing the lane's column
  unsigned long long mle = 0, mlh = 0;
  if (SRC) {
    for (int q = 0; q < np0; ++q)
      if (face_column(lf[q], i, j)) mle |= 1ull << q;
    for (int q = 0; q < np1; ++q)
      if (face_column(lf[np0 + q], i, j)) mlh |= 1ull << q;
  }
  const bool ndxy_e = SRC && step_nd_xy(sd, 0, x0, j), ndxy_h = SRC && step_nd_xy(sd, 1, x0, j);
  const bool ade_e = ADE && gi.ade[0].has, ade_h = ADE && gi.ade[1].has;
  unsigned axe[3] = {~0u, ~0u, ~0u}, axh[3] = {~0u, ~0u, ~0u};
  if (ade_e) pole_xy(gi.ade[0], i, j, axe);
  if (ade_h) pole_xy(gi.ade[1], i, j, axh);

  // H of step n on plane k0 - 1: the E update's z - 1 neighbours
  float hx_km = ld_at(plane_base(gi.f[3], (long long)k0 * pxy), ob);
  float hy_km = ld_at(plane_base(gi.f[4], (long long)k0 * pxy), ob);
  // the lane's constants arrive before the march: a load still in flight at
  // the loop head would make every plane's first use of it wait for all
  // memory operations issued since, stores included
  __builtin_amdgcn_s_waitcnt(0);
  // carried from plane k - 1 to the H update of that plane
  float hx_c = 0.0f, hy_c = 0.0f, hz_c = 0.0f;          // H of step n
  float ex_c = 0.0f, ey_c = 0.0f, ez_c = 0.0f;          // E of step n+1
  float ey_ip_c = 0.0f, ez_ip_c = 0.0f, ex_jp_c = 0.0f, ez_jp_c = 0.0f;
  // H coefficients {GM1, GM2} and dispersive slots of plane k - 1, looked up
  // while the plane's E half runs (the media loads never outlive the plane)
  float2 chx = make_float2(0.f, 0.f), chy = chx, chz = chx;
  int shx = -1, shy = -1, shz = -1;

  // One wait per plane: every load of the plane (fields, media, and the psi
  // and pole state of both halves) is issued first; every lane then computes
  // all components, lanes outside a component's sweep producing values that
  // no owned cell reads, and only the stores are masked.  A result that a
  // masked lane left as a loaded register would make the H half wait for the
  // E stores to drain (loads and stores share one counter).
  for (int k = k0; k <= k1; ++k) {
    const long long P = (long long)(k + 1) * pxy;     // plane k
    const int kh = k - 1;
    const bool hk = k > k0;
    const long long Ph = P - pxy;
    const bool eplane = k <= nz;
    CStepAux *a = per_plane(ax);
    // ---- loads of plane k.  (Issuing plane k + 1's loads before plane k's
    // arithmetic was measured: no faster, 16 more registers.)
    PlaneLoad cur{};
    if (eplane) {
      // media from the int16 arrays or their 8-bit copies (MT)
      if constexpr (sizeof(MT) == 1) {
        const unsigned om8 = ob >> 2;
        if constexpr (CPML || ADE || SRC)
          load_plane(cur, a->fi, a->mi8, P, ob, obx, oby, om8);
        else
          load_plane(cur, gi.f, sd.mi8, P, ob, obx, oby, om8);
      } else {
        if constexpr (CPML || ADE || SRC)
          load_plane(cur, a->fi, a->mi, P, ob, obx, oby, om);
        else
          load_plane(cur, gi.f, gi.mi, P, ob, obx, oby, om);
      }
    }
    const float hx0 = cur.hx, hy0 = cur.hy, hz0 = cur.hz, hz_jm = cur.hz_jm, hx_jm = cur.hx_jm,
                hz_im = cur.hz_im, hy_im = cur.hy_im, ex0 = cur.ex, ey0 = cur.ey, ez0 = cur.ez;
    const int mx = cur.mx, my = cur.my, mz = cur.mz, mhx = cur.mhx, mhy = cur.mhy,
              mhz = cur.mhz;
    float2 nhx = make_float2(0.f, 0.f), nhy = nhx, nhz = nhx;
    int nsx = -1, nsy = -1, nsz = -1;
    const float idzh = eplane ? ro_uniform(idzh_t, k + 1) : 0.0f;
    const bool ez_on = in && k < nz;
    const bool wr = own && k < k1;
    // CPML psi of the E corrections, in use order: Ex (y, z), Ey (z, x), Ez (y, x)
    bool f0 = false, f1 = false, f2 = false, f3 = false, f4 = false, f5 = false;
    float pv0 = 0.f, pv1 = 0.f, pv2 = 0.f, pv3 = 0.f, pv4 = 0.f, pv5 = 0.f;
    float bez = 0.f, cez = 0.f;
    int qzn = 0;
    if (CPML && eplane) {
      const bool szn = slab_node(k, nz, Lz0, Lz1, qzn);
      f0 = ex_on && syn; f1 = ex_on && szn;
      f2 = ey_on && szn; f3 = ey_on && sxn;
      f4 = ez_on && syn; f5 = ez_on && sxn;
      if (szn) { bez = ro_uniform(a->pbe[2], k + 1); cez = ro_uniform(a->pce[2], k + 1); }
      if (f0) pv0 = (a->psi_i[PSI_EXY] + k * Ly * nx)[(unsigned)(qyn * nx + i)];
      if (f1) pv1 = (a->psi_i[PSI_EXZ] + (long long)qzn * (ny + 1) * nx)[(unsigned)(j * nx + i)];
      if (f2) pv2 = (a->psi_i[PSI_EYZ] + (long long)qzn * ny * (nx + 1))[(unsigned)(j * (nx + 1) + i)];
      if (f3) pv3 = (a->psi_i[PSI_EYX] + k * ny * Lx)[(unsigned)(j * Lx + qxn)];
      if (f4) pv4 = (a->psi_i[PSI_EZY] + k * Ly * (nx + 1))[(unsigned)(qyn * (nx + 1) + i)];
      if (f5) pv5 = (a->psi_i[PSI_EZX] + k * (ny + 1) * Lx)[(unsigned)(j * Lx + qxn)];
    }
    // pole state of this plane's dispersive E edges
    bool g0 = false, g1 = false, g2 = false;
    long long z0 = 0, z1 = 0, z2 = 0;
    float ep0 = 0.f, ep1 = 0.f, ep2 = 0.f;
    jstate_t aj0[kPre], aj1[kPre], aj2[kPre];
    // the stored previous field is written unless it is left to the settling
    // pass (tiles without sources, whose next step takes it from the field)
    bool wrp_e = true;
    if (ade_e && eplane) {
      const auto &A = a->ai[0];
      g0 = pole_plane(A, 0, k, z0) && ex_on && axe[0] != ~0u;
      g1 = pole_plane(A, 1, k, z1) && ey_on && axe[1] != ~0u;
      g2 = pole_plane(A, 2, k, z2) && ez_on && axe[2] != ~0u;
      // the field of the previous step is the stored one unless nothing acted
      // on the edge after last step's pole update (no sources in the tile and
      // no fields set from outside since): then it is the field just loaded
      const bool pf = !SRC && sd.prev_from_field;
      wrp_e = !(pf && sd.lazy_prev);
      if (g0) { pole_load(A, 0, z0, axe[0], ep0, aj0, !pf); if (pf) ep0 = ex0; }
      if (g1) { pole_load(A, 1, z1, axe[1], ep1, aj1, !pf); if (pf) ep1 = ey0; }
      if (g2) { pole_load(A, 2, z2, axe[2], ep2, aj2, !pf); if (pf) ep2 = ez0; }
    }
    // the same for the H cells of plane k - 1: Hx (y, z), Hy (z, x), Hz (y, x)
    const bool hxk = hk && hx_on && kh < nz, hyk = hk && hy_on && kh < nz, hzk = hk && hz_on;
    bool r0 = false, r1 = false, r2 = false, r3 = false, r4 = false, r5 = false;
    float qv0 = 0.f, qv1 = 0.f, qv2 = 0.f, qv3 = 0.f, qv4 = 0.f, qv5 = 0.f;
    float bmz = 0.f, cmz = 0.f;
    int qzc = 0;
    if (CPML && hk) {
      const bool szc = slab_cell(kh, nz, Lz0, Lz1, qzc);
      r0 = hxk && syc; r1 = hxk && szc;
      r2 = hyk && szc; r3 = hyk && sxc;
      r4 = hzk && syc; r5 = hzk && sxc;
      if (szc) { bmz = ro_uniform(a->pbm[2], kh + 1); cmz = ro_uniform(a->pcm[2], kh + 1); }
      if (r0) qv0 = (a->psi_i[PSI_HXY] + kh * Ly * (nx + 1))[(unsigned)(qyc * (nx + 1) + i)];
      if (r1) qv1 = (a->psi_i[PSI_HXZ] + (long long)qzc * ny * (nx + 1))[(unsigned)(j * (nx + 1) + i)];
      if (r2) qv2 = (a->psi_i[PSI_HYZ] + (long long)qzc * (ny + 1) * nx)[(unsigned)(j * nx + i)];
      if (r3) qv3 = (a->psi_i[PSI_HYX] + kh * (ny + 1) * Lx)[(unsigned)(j * Lx + qxc)];
      if (r4) qv4 = (a->psi_i[PSI_HZY] + kh * Ly * nx)[(unsigned)(qyc * nx + i)];
      if (r5) qv5 = (a->psi_i[PSI_HZX] + kh * ny * Lx)[(unsigned)(j * Lx + qxc)];
    }
    // (the pole state of H cells, rare, is read where it is used: prefetched
    // here it would cost the pole classes 15 registers and spills)
    bool u0 = false, u1 = false, u2 = false;
    long long w0 = 0, w1 = 0, w2 = 0;
    const bool hpf = !SRC && sd.prev_from_field;
    const bool wrp_h = !(hpf && sd.lazy_prev);
    if (ade_h && hk) {
      const auto &A = a->ai[1];
      u0 = pole_plane(A, 0, kh, w0) && hxk && axh[0] != ~0u;
      u1 = pole_plane(A, 1, kh, w1) && hyk && axh[1] != ~0u;
      u2 = pole_plane(A, 2, kh, w2) && hzk && axh[2] != ~0u;
    }
    const float idze = hk ? ro_uniform(idze_t, kh + 1) : 0.0f;
    // TF/SF faces: the faces acting on each component of this lane's E cell
    // (plane k) and H cell (plane k - 1), and the first one's incident value
    unsigned long long be0 = 0, be1 = 0, be2 = 0, bh0 = 0, bh1 = 0, bh2 = 0;
    float ie0 = 0.f, ie1 = 0.f, ie2 = 0.f, ih0 = 0.f, ih1 = 0.f, ih2 = 0.f;
    if (SRC && lzm) {
      if (eplane) {
        const unsigned long long z = lzm[k - k0] & mle;
        if (z) {
          be0 = ex_on ? z & lcm[0] : 0; be1 = ey_on ? z & lcm[1] : 0;
          be2 = ez_on ? z & lcm[2] : 0;
          if (be0) ie0 = face_inc(sd.inc, lf[__builtin_ctzll(be0)], i, j, k);
          if (be1) ie1 = face_inc(sd.inc, lf[__builtin_ctzll(be1)], i, j, k);
          if (be2) ie2 = face_inc(sd.inc, lf[__builtin_ctzll(be2)], i, j, k);
        }
      }
      if (hk) {
        const unsigned long long z = lzm[kc1 + (kh - k0)] & mlh;
        if (z) {
          bh0 = hxk ? z & lcm[3] : 0; bh1 = hyk ? z & lcm[4] : 0; bh2 = hzk ? z & lcm[5] : 0;
          if (bh0) ih0 = face_inc(sd.inc, lf[np0 + __builtin_ctzll(bh0)], i, j, kh);
          if (bh1) ih1 = face_inc(sd.inc, lf[np0 + __builtin_ctzll(bh1)], i, j, kh);
          if (bh2) ih2 = face_inc(sd.inc, lf[np0 + __builtin_ctzll(bh2)], i, j, kh);
        }
      }
    }

    // ---- E of step n+1 on plane k (Advance_Ex/Ey/Ez, AdvanceelectricCPML,
    // AdvanceEDispersiveE, AdvancePlaneWaveE, AdvanceNodalE)
    float ex_n = 0.0f, ey_n = 0.0f, ez_n = 0.0f;
    if (eplane) {
      nhx = th.gg[mhx]; nhy = th.gg[mhy]; nhz = th.gg[mhz];
      if (ade_h) { nsx = th.slot[mhx]; nsy = th.slot[mhy]; nsz = th.slot[mhz]; }
      const bool nd = SRC && ndxy_e && k >= sd.cellbox[0][2] && k <= sd.cellbox[0][5];
      {
        const float2 c = te.gg[mx];
        float e = yee_upd(c.x, ex0, c.y, curl2(hz0 - hz_jm, idyh, hy0 - hy_km, idzh));
        if (CPML) {
          if (f0)
            e = madd(e, c.y, psi_new(pv0, (a->psi_o[PSI_EXY] + k * Ly * nx) + (unsigned)(qyn * nx + i),
                                     wr, bey, hz0 - hz_jm, cey));
          if (f1)
            e = madd(e, -c.y, psi_new(pv1, (a->psi_o[PSI_EXZ] + (long long)qzn * (ny + 1) * nx) +
                                               (unsigned)(j * nx + i),
                                      wr, bez, hy0 - hy_km, cez));
        }
        if (g0) {
          const int s = te.slot[mx];
          if (s >= 0) e = pole_new(a->ai[0], a->ao[0], 0, z0, axe[0], wr, wrp_e, a->dt, te, s, ep0, aj0, e);
        }
        if (SRC) {
          if (be0) e = faces_apply(lf, be0, ie0, sd.inc, sd.pwG[0], i, j, k, e);
          if (nd && ex_on)
            e = step_nodal(sd.src, 0, 0, P + (jc + 1) * (long long)gi.px + gi.xoff + ic, e, n);
        }
        ex_n = e;
        if (wr && ex_on) st_at(plane_base(fo_at(a, go, 0), P), ob, e);
      }
      {
        const float2 c = te.gg[my];
        float e = yee_upd(c.x, ey0, c.y, curl2(hx0 - hx_km, idzh, hz0 - hz_im, idxh));
        if (CPML) {
          if (f2)
            e = madd(e, c.y, psi_new(pv2, (a->psi_o[PSI_EYZ] + (long long)qzn * ny * (nx + 1)) +
                                              (unsigned)(j * (nx + 1) + i),
                                     wr, bez, hx0 - hx_km, cez));
          if (f3)
            e = madd(e, -c.y, psi_new(pv3, (a->psi_o[PSI_EYX] + k * ny * Lx) + (unsigned)(j * Lx + qxn),
                                      wr, bex, hz0 - hz_im, cex));
        }
        if (g1) {
          const int s = te.slot[my];
          if (s >= 0) e = pole_new(a->ai[0], a->ao[0], 1, z1, axe[1], wr, wrp_e, a->dt, te, s, ep1, aj1, e);
        }
        if (SRC) {
          if (be1) e = faces_apply(lf, be1, ie1, sd.inc, sd.pwG[0], i, j, k, e);
          if (nd && ey_on)
            e = step_nodal(sd.src, 0, 1, P + (jc + 1) * (long long)gi.px + gi.xoff + ic, e, n);
        }
        ey_n = e;
        if (wr && ey_on) st_at(plane_base(fo_at(a, go, 1), P), ob, e);
      }
      {
        const float2 c = te.gg[mz];
        float e = yee_upd(c.x, ez0, c.y, curl2(hy0 - hy_im, idxh, hx0 - hx_jm, idyh));
        if (CPML) {
          if (f4)
            e = madd(e, -c.y, psi_new(pv4, (a->psi_o[PSI_EZY] + k * Ly * (nx + 1)) +
                                               (unsigned)(qyn * (nx + 1) + i),
                                      wr, bey, hx0 - hx_jm, cey));
          if (f5)
            e = madd(e, c.y, psi_new(pv5, (a->psi_o[PSI_EZX] + k * (ny + 1) * Lx) + (unsigned)(j * Lx + qxn),
                                     wr, bex, hy0 - hy_im, cex));
        }
        if (g2) {
          const int s = te.slot[mz];
          if (s >= 0) e = pole_new(a->ai[0], a->ao[0], 2, z2, axe[2], wr, wrp_e, a->dt, te, s, ep2, aj2, e);
        }
        if (SRC) {
          if (be2) e = faces_apply(lf, be2, ie2, sd.inc, sd.pwG[0], i, j, k, e);
          if (nd && ez_on)
            e = step_nodal(sd.src, 0, 2, P + (jc + 1) * (long long)gi.px + gi.xoff + ic, e, n);
        }
        ez_n = e;
        if (wr && ez_on) st_at(plane_base(fo_at(a, go, 2), P), ob, e);
      }
    }
    // ---- E of plane k at x + 1 (next lane) and y + 1 (next row).  (With
    // half-width rows the last lane of a row receives the next row's first
    // value: that lane is at most the halo column, whose H is not stored.)
    const float ey_ip_n = __shfl_down(ey_n, 1), ez_ip_n = __shfl_down(ez_n, 1);
    const int par = (k - k0) & 1;
    xch[(par * RR + row) * XW + lane] = make_float2(ex_n, ez_n);
    __syncthreads();
    float ex_jp_n = 0.0f, ez_jp_n = 0.0f;
    if (row + 1 < RR) {
      const float2 v = xch[(par * RR + row + 1) * XW + lane];
      ex_jp_n = v.x;
      ez_jp_n = v.y;
    }
    // ---- H of step n+1/2 on plane k - 1 (Advance_Hx/Hy/Hz,
    // AdvanceMagneticCPML, AdvanceMdispersiveH, AdvancePlaneWaveH,
    // AdvanceNodalH)
    if (hk) {
      const bool nd = SRC && ndxy_h && kh >= sd.cellbox[1][2] && kh <= sd.cellbox[1][5];
      {
        const float2 c = chx;
        float h = yee_upd(c.x, hx_c, c.y, curl2(ey_n - ey_c, idze, ez_jp_c - ez_c, idye));
        if (CPML) {
          if (r0)
            h = madd(h, -c.y, psi_new(qv0, (a->psi_o[PSI_HXY] + kh * Ly * (nx + 1)) +
                                               (unsigned)(qyc * (nx + 1) + i),
                                      true, bmy, ez_jp_c - ez_c, cmy));
          if (r1)
            h = madd(h, c.y, psi_new(qv1, (a->psi_o[PSI_HXZ] + (long long)qzc * ny * (nx + 1)) +
                                              (unsigned)(j * (nx + 1) + i),
                                     true, bmz, ey_n - ey_c, cmz));
        }
        if (u0) {
          const int s = shx;
          if (s >= 0) {
            float hp = hx_c;
            jstate_t bj[kPre];
            pole_load(a->ai[1], 0, w0, axh[0], hp, bj, !hpf);
            h = pole_new(a->ai[1], a->ao[1], 0, w0, axh[0], true, wrp_h, a->dt, th, s, hp, bj, h);
          }
        }
        if (SRC) {
          if (bh0) h = faces_apply(lf + np0, bh0, ih0, sd.inc, sd.pwG[1], i, j, kh, h);
          if (nd && hxk)
            h = step_nodal(sd.src, 1, 3, Ph + (jc + 1) * (long long)gi.px + gi.xoff + ic, h, n);
        }
        if (hxk) st_at(plane_base(fo_at(a, go, 3), Ph), ob, h);
      }
      {
        const float2 c = chy;
        float h = yee_upd(c.x, hy_c, c.y, curl2(ez_ip_c - ez_c, idxe, ex_n - ex_c, idze));
        if (CPML) {
          if (r2)
            h = madd(h, -c.y, psi_new(qv2, (a->psi_o[PSI_HYZ] + (long long)qzc * (ny + 1) * nx) +
                                               (unsigned)(j * nx + i),
                                      true, bmz, ex_n - ex_c, cmz));
          if (r3)
            h = madd(h, c.y, psi_new(qv3, (a->psi_o[PSI_HYX] + kh * (ny + 1) * Lx) +
                                              (unsigned)(j * Lx + qxc),
                                     true, bmx, ez_ip_c - ez_c, cmx));
        }
        if (u1) {
          const int s = shy;
          if (s >= 0) {
            float hp = hy_c;
            jstate_t bj[kPre];
            pole_load(a->ai[1], 1, w1, axh[1], hp, bj, !hpf);
            h = pole_new(a->ai[1], a->ao[1], 1, w1, axh[1], true, wrp_h, a->dt, th, s, hp, bj, h);
          }
        }
        if (SRC) {
          if (bh1) h = faces_apply(lf + np0, bh1, ih1, sd.inc, sd.pwG[1], i, j, kh, h);
          if (nd && hyk)
            h = step_nodal(sd.src, 1, 4, Ph + (jc + 1) * (long long)gi.px + gi.xoff + ic, h, n);
        }
        if (hyk) st_at(plane_base(fo_at(a, go, 4), Ph), ob, h);
      }
      {
        const float2 c = chz;
        float h = yee_upd(c.x, hz_c, c.y, curl2(ex_jp_c - ex_c, idye, ey_ip_c - ey_c, idxe));
        if (CPML) {
          if (r4)
            h = madd(h, c.y, psi_new(qv4, (a->psi_o[PSI_HZY] + kh * Ly * nx) + (unsigned)(qyc * nx + i),
                                     true, bmy, ex_jp_c - ex_c, cmy));
          if (r5)
            h = madd(h, -c.y, psi_new(qv5, (a->psi_o[PSI_HZX] + kh * ny * Lx) + (unsigned)(j * Lx + qxc),
                                      true, bmx, ey_ip_c - ey_c, cmx));
        }
        if (u2) {
          const int s = shz;
          if (s >= 0) {
            float hp = hz_c;
            jstate_t bj[kPre];
            pole_load(a->ai[1], 2, w2, axh[2], hp, bj, !hpf);
            h = pole_new(a->ai[1], a->ao[1], 2, w2, axh[2], true, wrp_h, a->dt, th, s, hp, bj, h);
          }
        }
        if (SRC) {
          if (bh2) h = faces_apply(lf + np0, bh2, ih2, sd.inc, sd.pwG[1], i, j, kh, h);
          if (nd && hzk)
            h = step_nodal(sd.src, 1, 5, Ph + (jc + 1) * (long long)gi.px + gi.xoff + ic, h, n);
        }
        if (hzk) st_at(plane_base(fo_at(a, go, 5), Ph), ob, h);
      }
    }
    // ---- carry plane k to the next iteration
    hx_km = hx0; hy_km = hy0;
    hx_c = hx0; hy_c = hy0; hz_c = hz0;
    ex_c = ex_n; ey_c = ey_n; ez_c = ez_n;
    ey_ip_c = ey_ip_n; ez_ip_c = ez_ip_n; ex_jp_c = ex_jp_n; ez_jp_c = ez_jp_n;
    chx = nhx; chy = nhy; chz = nhz;
    shx = nsx; shy = nsy; shz = nsz;
  }
}

template <int R, bool CPML, bool ADE, bool SRC, class MT, int XW>
__global__ void __launch_bounds__(64 * R) k_step(StepDev sd) {
  step_tile<R, CPML, ADE, SRC, MT, XW>(sd, sd.tiles[xcd_remap(blockIdx.x, sd.tm.nblk)]);
}

// The same tiles of many grids in one launch (batches of grids): block b runs
// tile bt[b].t of grid bt[b].grid, whose descriptor is sds[bt[b].grid].
template <int R, bool CPML, bool ADE, bool SRC, class MT, int XW>
__global__ void __launch_bounds__(64 * R) k_step_many(const StepDev *__restrict__ sds,
                                                      const GridTile *__restrict__ bt, int n) {
  const GridTile g = bt[xcd_remap(blockIdx.x, n)];
  step_tile<R, CPML, ADE, SRC, MT, XW>(sds[g.grid], g.t);
}

// ---- plain interior tiles, four x cells per lane ----------------------------
// k_step4 is k_step's plain path (no CPML slab, pole state or source in the
// tile) with each lane owning four consecutive x cells: one 16-byte load or
// store per array, plane and lane, a quarter of the memory instructions per
// cell.  A tile starts on a 16-byte boundary (x0 % 4 == 0) and owns a
// multiple of four columns (wx <= 252), so lane wx / 4 holds the halo column;
// every owned column and the halo lie inside the x sweep of all six
// components (x0 + wx + 1 <= nx), so a lane's four cells are stored together
// or not at all.  Inside a lane the x neighbours are the lane's own cells;
// across lanes the E values at x + 1 and the H values at x - 1 move by DPP
// wave shifts, lane 0's H at x0 - 1 is a scalar load.  Rows exchange E
// through LDS as in k_step.  Per cell the arithmetic and its order are
// k_step's, so the results are the same bits.

// 16-byte and 8-byte vectors as the compiler's own vector types (these load
// and store through global-address-space pointers)
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ f32x4 v4(float a, float b, float c, float d) {
  f32x4 r;
  r.x = a; r.y = b; r.z = c; r.w = d;
  return r;
}

// lane l receives lane l - 1's v (wave_shr:1); lane 0 receives first
__device__ __forceinline__ float from_left(float v, float first) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, first),
                                                               __builtin_bit_cast(int, v), 0x138,
                                                               0xf, 0xf, false));
}
// lane l receives lane l + 1's v (wave_shl:1); lane 63 receives 0
__device__ __forceinline__ float from_right(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                               0x130, 0xf, 0xf, false));
}

// the four media indices of a lane's cells, from the 8-bit copies (one
// dword) or the int16 arrays (two dwords)
template <class MT> struct Media4;
template <> struct Media4<uint8_t> {
  typedef uint32_t raw;
  static __device__ __forceinline__ raw ld(const uint8_t *base, long long P, unsigned ob) {
    return ld_at(reinterpret_cast<const SFD_GLOBAL uint32_t *>(plane_base(base, P)), ob >> 2);
  }
  static __device__ __forceinline__ int at(raw r, int c) { return (int)((r >> (8 * c)) & 0xffu); }
};
template <> struct Media4<int16_t> {
  typedef u32x2 raw;
  static __device__ __forceinline__ raw ld(const int16_t *base, long long P, unsigned ob) {
    return ld_at(reinterpret_cast<const SFD_GLOBAL u32x2 *>(plane_base(base, P)), ob >> 1);
  }
  static __device__ __forceinline__ int at(raw r, int c) {
    return (int)(((c < 2 ? r.x : r.y) >> (16 * (c & 1))) & 0xffffu);
  }
};

// a wave-uniform field value (scalar load: read-only data within the launch)
__device__ __forceinline__ float ld_uniform(const float *base, long long P, unsigned ob) {
  return ((const __attribute__((address_space(4))) float *)(base + P))[ob >> 2];
}

static __host__ __device__ __forceinline__ size_t step4_xch_offset(int nmedia, int nd0, int nd1) {
  return (step_tab_bytes(nmedia, nd0) + step_tab_bytes(nmedia, nd1) + 15) & ~(size_t)15;
}

// UNI: every cell the tile reads holds one medium (carried in k1's high
// bits): no media loads, one coefficient pair per field for the whole tile
template <int R, class MT, bool UNI = false>
__device__ __forceinline__ void step4_tile(const StepDev &sd, const StepTile tl) {
  typedef Media4<MT> M4;
  typedef typename M4::raw raw;
  extern __shared__ float2 lds[];
  const DevGrid &gi = sd.gi;
  const DevGrid &go = sd.go;
  const size_t offh = step_tab_bytes(gi.nmedia, gi.ade[0].ndisp);
  const size_t offx = step4_xch_offset(gi.nmedia, gi.ade[0].ndisp, gi.ade[1].ndisp);
  const MedTab te = stage_tables<true, false>(gi, gi.g1, gi.g2, lds);
  const MedTab th = stage_tables<true, false>(
      gi, gi.gm1, gi.gm2, reinterpret_cast<float2 *>(reinterpret_cast<char *>(lds) + offh));
  // E exchange between rows: [2 parities][R rows][64 lanes] of Ex, then of Ez
  f32x4 *xe = reinterpret_cast<f32x4 *>(reinterpret_cast<char *>(lds) + offx);
  f32x4 *xz = xe + 2 * R * kWave;
  const int x0 = tl.x & 0xfffff, wx = tl.x >> 20;
  const int y0 = tl.y & 0xfffff, wy = tl.y >> 20;
  const int k0 = tl.k0;
  const int lane = threadIdx.x;
  const int row = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
  const int nx = gi.nx, ny = gi.ny, nz = gi.nz;
  // the lane's first cell; lanes and rows past the halo repeat the halo's
  // (their results are never used, their loads hit the halo's lines)
  const int j = y0 + min(row, wy);
  const int i = x0 + 4 * min(lane, wx >> 2);
  const bool own = 4 * lane < wx && row < wy && j <= ny;
  const int k1 = min((int)((unsigned)tl.k1 & kTileK1Mask), nz + 1);
  const int m0 = UNI ? (int)((unsigned)tl.k1 >> kTileMediumShift) : 0;
  // lanes past the grid's edge address the last whole vector of the row
  // (their results are never used): every access stays inside its array
  const int ic = min(i, (nx + 3) & ~3), jc = min(j, ny);
  const long long pxy = gi.pxy;
  const unsigned pxb = (unsigned)gi.px * 4u;
  const unsigned ob = (unsigned)(jc + 1) * pxb + (unsigned)(gi.xoff + ic) * 4u;
  const unsigned oby = ob - pxb;
  const unsigned oe = (unsigned)(jc + 1) * pxb + (unsigned)(gi.xoff + x0 - 1) * 4u;  // x0 - 1
  float idxh[4], idxe[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int q = min(ic + c, nx) + 1;
    idxh[c] = gi.id[3][q];
    idxe[c] = gi.id[0][q];
  }
  const float idyh = ro_uniform(gi.id[4], jc + 1), idye = ro_uniform(gi.id[1], jc + 1);
  const float *idzh_t = gi.id[5], *idze_t = gi.id[2];
  // H of step n on plane k0 - 1: the E update's z - 1 neighbours
  f32x4 hx_c = ld_at(reinterpret_cast<const SFD_GLOBAL f32x4 *>(
                          plane_base(gi.f[3], (long long)k0 * pxy)), ob);
  f32x4 hy_c = ld_at(reinterpret_cast<const SFD_GLOBAL f32x4 *>(
                          plane_base(gi.f[4], (long long)k0 * pxy)), ob);
  __builtin_amdgcn_s_waitcnt(0);
  f32x4 hz_c = v4(0.f, 0.f, 0.f, 0.f);
  f32x4 ex_c = hz_c, ey_c = hz_c, ez_c = hz_c, exj_c = hz_c, ezj_c = hz_c;
  float ey3_c = 0.f, ez3_c = 0.f;      // Ey, Ez at the next lane's first cell
  raw mhx_c{}, mhy_c{}, mhz_c{};       // H media of plane k - 1
  for (int k = k0; k <= k1; ++k) {
    const long long P = (long long)(k + 1) * pxy;
    const bool eplane = k <= nz;
    const bool hk = k > k0;
    const int kh = k - 1;
    f32x4 hx = v4(0.f, 0.f, 0.f, 0.f), hy = hx, hz = hx, hzj = hx, hxj = hx;
    f32x4 ex = hx, ey = hx, ez = hx;
    raw mx{}, my{}, mz{}, mhx{}, mhy{}, mhz{};
    float hze = 0.f, hye = 0.f;
    if (eplane) {
      const SFD_GLOBAL f32x4 *Hx = reinterpret_cast<const SFD_GLOBAL f32x4 *>(plane_base(gi.f[3], P));
      const SFD_GLOBAL f32x4 *Hy = reinterpret_cast<const SFD_GLOBAL f32x4 *>(plane_base(gi.f[4], P));
      const SFD_GLOBAL f32x4 *Hz = reinterpret_cast<const SFD_GLOBAL f32x4 *>(plane_base(gi.f[5], P));
      hx = ld_at(Hx, ob); hy = ld_at(Hy, ob); hz = ld_at(Hz, ob);
      hzj = ld_at(Hz, oby); hxj = ld_at(Hx, oby);
      ex = ld_at(reinterpret_cast<const SFD_GLOBAL f32x4 *>(plane_base(gi.f[0], P)), ob);
      ey = ld_at(reinterpret_cast<const SFD_GLOBAL f32x4 *>(plane_base(gi.f[1], P)), ob);
      ez = ld_at(reinterpret_cast<const SFD_GLOBAL f32x4 *>(plane_base(gi.f[2], P)), ob);
      if constexpr (!UNI) {
        const MT *mi[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          if constexpr (sizeof(MT) == 1) mi[c] = sd.mi8[c];
          else mi[c] = gi.mi[c];
        }
        mx = M4::ld(mi[0], P, ob); my = M4::ld(mi[1], P, ob); mz = M4::ld(mi[2], P, ob);
        mhx = M4::ld(mi[3], P, ob); mhy = M4::ld(mi[4], P, ob); mhz = M4::ld(mi[5], P, ob);
      }
      hze = ld_uniform(gi.f[5], P, oe);
      hye = ld_uniform(gi.f[4], P, oe);
    }
    const float idzh = eplane ? ro_uniform(idzh_t, k + 1) : 0.0f;
    const float idze = hk ? ro_uniform(idze_t, kh + 1) : 0.0f;
    const bool wr = own && k < k1;
    // ---- E of step n+1 on plane k
    f32x4 exn = v4(0.f, 0.f, 0.f, 0.f), eyn = exn, ezn = exn;
    if (eplane) {
      const float hzm[4] = {from_left(hz.w, hze), hz.x, hz.y, hz.z};   // Hz at x - 1
      const float hym[4] = {from_left(hy.w, hye), hy.x, hy.y, hy.z};   // Hy at x - 1
      const float HX[4] = {hx.x, hx.y, hx.z, hx.w}, HY[4] = {hy.x, hy.y, hy.z, hy.w},
                  HZ[4] = {hz.x, hz.y, hz.z, hz.w};
      const float HZJ[4] = {hzj.x, hzj.y, hzj.z, hzj.w}, HXJ[4] = {hxj.x, hxj.y, hxj.z, hxj.w};
      const float HXK[4] = {hx_c.x, hx_c.y, hx_c.z, hx_c.w}, HYK[4] = {hy_c.x, hy_c.y, hy_c.z, hy_c.w};
      const float EX[4] = {ex.x, ex.y, ex.z, ex.w}, EY[4] = {ey.x, ey.y, ey.z, ey.w},
                  EZ[4] = {ez.x, ez.y, ez.z, ez.w};
      float rx[4], ry[4], rz[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const float2 cx = te.gg[UNI ? m0 : M4::at(mx, c)], cy = te.gg[UNI ? m0 : M4::at(my, c)],
                     cz = te.gg[UNI ? m0 : M4::at(mz, c)];
        rx[c] = yee_upd(cx.x, EX[c], cx.y, curl2(HZ[c] - HZJ[c], idyh, HY[c] - HYK[c], idzh));
        ry[c] = yee_upd(cy.x, EY[c], cy.y, curl2(HX[c] - HXK[c], idzh, HZ[c] - hzm[c], idxh[c]));
        rz[c] = yee_upd(cz.x, EZ[c], cz.y, curl2(HY[c] - hym[c], idxh[c], HX[c] - HXJ[c], idyh));
      }
      exn = v4(rx[0], rx[1], rx[2], rx[3]);
      eyn = v4(ry[0], ry[1], ry[2], ry[3]);
      ezn = v4(rz[0], rz[1], rz[2], rz[3]);
      if (wr) st_at(reinterpret_cast<SFD_GLOBAL f32x4 *>(plane_base(go.f[0], P)), ob, exn);
      if (wr && j < ny) st_at(reinterpret_cast<SFD_GLOBAL f32x4 *>(plane_base(go.f[1], P)), ob, eyn);
      if (wr && k < nz) st_at(reinterpret_cast<SFD_GLOBAL f32x4 *>(plane_base(go.f[2], P)), ob, ezn);
    }
    // ---- E of plane k at x + 1 (next lane) and y + 1 (next row)
    const float ey3 = from_right(eyn.x), ez3 = from_right(ezn.x);
    const int par = (k - k0) & 1;
    xe[(par * R + row) * kWave + lane] = exn;
    xz[(par * R + row) * kWave + lane] = ezn;
    __syncthreads();
    f32x4 exj = v4(0.f, 0.f, 0.f, 0.f), ezj = exj;
    if (row + 1 < R) {
      exj = xe[(par * R + row + 1) * kWave + lane];
      ezj = xz[(par * R + row + 1) * kWave + lane];
    }
    // ---- H of step n+1/2 on plane k - 1
    if (hk) {
      const float EXC[4] = {ex_c.x, ex_c.y, ex_c.z, ex_c.w}, EYC[4] = {ey_c.x, ey_c.y, ey_c.z, ey_c.w},
                  EZC[4] = {ez_c.x, ez_c.y, ez_c.z, ez_c.w};
      const float EXN[4] = {exn.x, exn.y, exn.z, exn.w}, EYN[4] = {eyn.x, eyn.y, eyn.z, eyn.w};
      const float EXJ[4] = {exj_c.x, exj_c.y, exj_c.z, exj_c.w}, EZJ[4] = {ezj_c.x, ezj_c.y, ezj_c.z, ezj_c.w};
      const float EYP[4] = {ey_c.y, ey_c.z, ey_c.w, ey3_c}, EZP[4] = {ez_c.y, ez_c.z, ez_c.w, ez3_c};
      const float HXC[4] = {hx_c.x, hx_c.y, hx_c.z, hx_c.w}, HYC[4] = {hy_c.x, hy_c.y, hy_c.z, hy_c.w},
                  HZC[4] = {hz_c.x, hz_c.y, hz_c.z, hz_c.w};
      float rx[4], ry[4], rz[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const float2 cx = th.gg[UNI ? m0 : M4::at(mhx_c, c)],
                     cy = th.gg[UNI ? m0 : M4::at(mhy_c, c)],
                     cz = th.gg[UNI ? m0 : M4::at(mhz_c, c)];
        rx[c] = yee_upd(cx.x, HXC[c], cx.y, curl2(EYN[c] - EYC[c], idze, EZJ[c] - EZC[c], idye));
        ry[c] = yee_upd(cy.x, HYC[c], cy.y, curl2(EZP[c] - EZC[c], idxe[c], EXN[c] - EXC[c], idze));
        rz[c] = yee_upd(cz.x, HZC[c], cz.y, curl2(EXJ[c] - EXC[c], idye, EYP[c] - EYC[c], idxe[c]));
      }
      const long long Ph = P - pxy;
      if (own && kh < nz && j < ny)
        st_at(reinterpret_cast<SFD_GLOBAL f32x4 *>(plane_base(go.f[3], Ph)), ob,
              v4(rx[0], rx[1], rx[2], rx[3]));
      if (own && kh < nz)
        st_at(reinterpret_cast<SFD_GLOBAL f32x4 *>(plane_base(go.f[4], Ph)), ob,
              v4(ry[0], ry[1], ry[2], ry[3]));
      if (own && j < ny)
        st_at(reinterpret_cast<SFD_GLOBAL f32x4 *>(plane_base(go.f[5], Ph)), ob,
              v4(rz[0], rz[1], rz[2], rz[3]));
    }
    // ---- carry plane k to the next iteration
    hx_c = hx; hy_c = hy; hz_c = hz;
    ex_c = exn; ey_c = eyn; ez_c = ezn;
    exj_c = exj; ezj_c = ezj;
    ey3_c = ey3; ez3_c = ez3;
    mhx_c = mhx; mhy_c = mhy; mhz_c = mhz;
  }
}

template <int R, class MT, bool UNI>
__global__ void __launch_bounds__(64 * R) k_step4(StepDev sd) {
  step4_tile<R, MT, UNI>(sd, sd.tiles[xcd_remap(blockIdx.x, sd.tm.nblk)]);
}

template <int R, class MT>
__global__ void __launch_bounds__(64 * R) k_step4_many(const StepDev *__restrict__ sds,
                                                       const GridTile *__restrict__ bt, int n) {
  const GridTile g = bt[xcd_remap(blockIdx.x, n)];
  step4_tile<R, MT>(sds[g.grid], g.t);
}

size_t step4_lds_bytes(const DevGrid &g, int rows) {
  return step4_xch_offset(g.nmedia, g.ade[0].ndisp, g.ade[1].ndisp) +
         (size_t)2 * 2 * rows * kWave * 16;
}

// The step kernels of one class over n blocks: the tiles of one grid
// (sds == nullptr: k_step / k_step4 on sd) or tiles of many grids (k_step_many
// / k_step4_many on sds and bt).  narrow: media read from the 8-bit copies.
// one grid: the nuni single-medium tiles that lead the class's tiles, then
// the rest; a batch's tiles all run the general kernel
static int launch_step4_any(const StepDev &sd, const StepDev *sds, const GridTile *bt, int n,
                            int narrow, int rows, size_t lds, hipStream_t s, int nuni = 0) {
  const dim3 grid((unsigned)n);
  if (n <= 0) return 0;
  nuni = sds ? 0 : std::min(std::max(nuni, 0), n);
  StepDev su = sd, sm = sd;
  su.tm.nblk = nuni;
  sm.tiles = sd.tiles + nuni;
  sm.tm.nblk = n - nuni;
#define SFD_STEP4_M(RR, M)                                                                    \
  do {                                                                                        \
    if (sds) {                                                                                \
      hipLaunchKernelGGL((k_step4_many<RR, M>), grid, dim3(kWave, RR), lds, s, sds, bt, n);  \
    } else {                                                                                  \
      if (nuni > 0)                                                                           \
        hipLaunchKernelGGL((k_step4<RR, M, true>), dim3((unsigned)nuni), dim3(kWave, RR), lds, \
                           s, su);                                                            \
      if (n > nuni)                                                                           \
        hipLaunchKernelGGL((k_step4<RR, M, false>), dim3((unsigned)(n - nuni)),              \
                           dim3(kWave, RR), lds, s, sm);                                      \
    }                                                                                         \
  } while (0)
#define SFD_STEP4(RR)                    \
  do {                                   \
    if (narrow) SFD_STEP4_M(RR, uint8_t); \
    else SFD_STEP4_M(RR, int16_t);       \
  } while (0)
  if (rows == 4) SFD_STEP4(4);
  else if (rows == 8) SFD_STEP4(8);
  else if (rows == 16) SFD_STEP4(16);
  else return (int)hipErrorInvalidValue;
#undef SFD_STEP4
#undef SFD_STEP4_M
  return (int)hipGetLastError();
}

int launch_step4(const StepDev &sd, int rows, void *stream) {
  return launch_step4_any(sd, nullptr, nullptr, sd.tm.nblk, sd.narrow, rows,
                          step4_lds_bytes(sd.gi, rows), (hipStream_t)stream);
}

size_t step_lds_bytes(const DevGrid &g, int rows, int npw, int kc) {
  const size_t x = step_tab_bytes(g.nmedia, g.ade[0].ndisp) +
                   step_tab_bytes(g.nmedia, g.ade[1].ndisp) +
                   (size_t)2 * rows * kWave * sizeof(float2);
  return npw ? step_lds_faces(x, npw, kc, nullptr, nullptr) : x;
}

// class cls: 0-7 the k_step variant (bits 1 CPML, 2 poles, 4 sources), 8 the
// four-cell kernel, 9-16 the half-width variant of class cls - 9
static int launch_step_any(const StepDev &sd, const StepDev *sds, const GridTile *bt, int n,
                           int narrow, int rows, int cls, size_t lds, hipStream_t s,
                           int nuni = 0) {
  if (cls == 8) return launch_step4_any(sd, sds, bt, n, narrow, rows, lds, s, nuni);
  const dim3 grid((unsigned)n);
  if (n <= 0) return 0;
  const bool half = cls >= 9;
  if (half) cls -= 9;
#define SFD_STEP_K(RR, C, A, S, M, W)                                                         \
  do {                                                                                        \
    if (sds)                                                                                  \
      hipLaunchKernelGGL((k_step_many<RR, C, A, S, M, W>), grid, dim3(kWave, RR), lds, s, sds, \
                         bt, n);                                                              \
    else                                                                                      \
      hipLaunchKernelGGL((k_step<RR, C, A, S, M, W>), grid, dim3(kWave, RR), lds, s, sd);    \
  } while (0)
#define SFD_STEP_W(RR, C, A, S, W)                        \
  do {                                                    \
    if (narrow) SFD_STEP_K(RR, C, A, S, uint8_t, W);      \
    else SFD_STEP_K(RR, C, A, S, int16_t, W);             \
  } while (0)
#define SFD_STEP(RR, C, A, S)                  \
  do {                                         \
    if (half) SFD_STEP_W(RR, C, A, S, 32);     \
    else SFD_STEP_W(RR, C, A, S, 64);          \
  } while (0)
#define SFD_STEP_R(RR)                                         \
  switch (cls & 7) {                                           \
    case 0: SFD_STEP(RR, false, false, false); break;          \
    case 1: SFD_STEP(RR, true, false, false); break;           \
    case 2: SFD_STEP(RR, false, true, false); break;           \
    case 3: SFD_STEP(RR, true, true, false); break;            \
    case 4: SFD_STEP(RR, false, false, true); break;           \
    case 5: SFD_STEP(RR, true, false, true); break;            \
    case 6: SFD_STEP(RR, false, true, true); break;            \
    default: SFD_STEP(RR, true, true, true); break;            \
  }
  if (rows == 4) {
    SFD_STEP_R(4)
  } else if (rows == 6) {
    SFD_STEP_R(6)
  } else if (rows == 8) {
    SFD_STEP_R(8)
  } else if (rows == 16) {
    SFD_STEP_R(16)
  } else {
    return (int)hipErrorInvalidValue;
  }
#undef SFD_STEP_R
#undef SFD_STEP
#undef SFD_STEP_W
#undef SFD_STEP_K
  return (int)hipGetLastError();
}

int launch_step(const StepDev &sd, int rows, int cls, int nuni, void *stream) {
  const size_t lds = cls == 8 ? step4_lds_bytes(sd.gi, rows)
                              : step_lds_bytes(sd.gi, rows, sd.npw[0] + sd.npw[1], sd.tm.kc);
  return launch_step_any(sd, nullptr, nullptr, sd.tm.nblk, sd.narrow, rows, cls, lds,
                         (hipStream_t)stream, nuni);
}

int launch_step_many(const StepDev *sds, const GridTile *tiles, int n, int narrow, int rows,
                     int cls, size_t lds, void *stream) {
  const StepDev none{};
  return launch_step_any(none, sds, tiles, n, narrow, rows, cls, lds, (hipStream_t)stream);
}

}  // namespace SFD_NS
}  // namespace sfd
