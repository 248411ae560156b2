// fdtd_plan.cpp -- the tile plan of the one-launch step (k_step, k_step4):
// which block owns which cells, and which kernel variant (tile class) runs
// it.  Pure host code: the CPU tests call it through sfd_plan_step without a
// device.
//
// A tile owns columns [x0, x0+wx), rows [y0, y0+wy) and planes [k0, k1) and
// also computes E on its halo column x0+wx, row y0+wy and plane k1 (the E
// neighbours of its last H cells).  Its class is the code it needs: bit 0
// CPML (it reaches a CPML slab), bit 1 poles (it meets the pole-state box of
// some component), bit 2 sources (it meets a TF/SF face or the box of the
// nodal cells); class 8 is a plain tile run four x cells per lane.  The class
// of a tile is judged on its cells including the halo, where the kernels
// compute E too.
//
// Each axis is cut where what a tile meets changes (the CPML slab edges and
// the planes just outside each face and box), narrow pieces are folded into
// a neighbour (a 64-lane block owning a few columns wastes the rest), and
// the pieces are split into tiles no larger than a block owns: 63 columns
// (or 252 in four-cell tiles, which start on a 16-byte boundary and own a
// multiple of four columns), R - 1 rows, kc planes.  Plain tiles stacked in z
// merge into one tile of up to 64 planes.  A piece of at most 31 columns (the
// x CPML slabs with the faces next to them) runs half-width tiles: 32 lanes
// per row, two rows per wave, up to 2E - 1 rows per tile for a class on E
// waves per block.
#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "fdtd_internal.h"
#include "fdtd_plan.h"

namespace sfd_host {
using namespace sfd;
namespace {

struct Seg {
  int lo, hi;  // [lo, hi)
};

// the pieces of [0, N) between sorted cut points, narrow ones folded into
// their narrower neighbour until none is below minw (or one piece is left)
std::vector<Seg> segments(std::vector<int> cuts, int N, int minw) {
  cuts.push_back(0);
  cuts.push_back(N);
  for (int &c : cuts) c = std::min(std::max(c, 0), N);
  std::sort(cuts.begin(), cuts.end());
  cuts.erase(std::unique(cuts.begin(), cuts.end()), cuts.end());
  std::vector<Seg> s;
  for (size_t q = 0; q + 1 < cuts.size(); ++q) s.push_back({cuts[q], cuts[q + 1]});
  while (s.size() > 1) {
    size_t best = s.size();
    for (size_t q = 0; q < s.size(); ++q)
      if (s[q].hi - s[q].lo < minw && (best == s.size() || s[q].hi - s[q].lo < s[best].hi - s[best].lo))
        best = q;
    if (best == s.size()) break;
    size_t other;
    if (best == 0) other = 1;
    else if (best + 1 == s.size()) other = best - 1;
    else {
      const int wl = s[best - 1].hi - s[best - 1].lo, wr = s[best + 1].hi - s[best + 1].lo;
      other = wl <= wr ? best - 1 : best + 1;
    }
    const size_t a = std::min(best, other), b = std::max(best, other);
    s[a].hi = s[b].hi;
    s.erase(s.begin() + (long)b);
  }
  return s;
}

// [lo, hi) split into ceil(w / W) near-equal pieces of `unit`-multiples
void split(int lo, int hi, int W, int unit, std::vector<Seg> *out) {
  const int units = (hi - lo) / unit;
  const int wu = W / unit;
  if (units <= 0) return;
  const int nt = (units + wu - 1) / wu;
  const int base = units / nt, rem = units % nt;
  int x = lo;
  for (int t = 0; t < nt; ++t) {
    const int w = (base + (t < rem ? 1 : 0)) * unit;
    out->push_back({x, x + w});
    x += w;
  }
}

struct XTile {
  int lo, hi;
  bool vec;   // a four-cell tile (plain) or its split into 63-column tiles
  bool half;  // at most 31 columns: half-width tiles
};

}  // namespace

int plan_classify(const PlanGeom &g, const int lo[3], const int hi[3]) {
  int cls = 0;
  for (int a = 0; a < 3; ++a)
    if ((g.L[a][0] > 0 && lo[a] < g.L[a][0]) || (g.L[a][1] > 0 && hi[a] >= g.n[a] - g.L[a][1]))
      cls |= 1;
  for (const PlanBox &b : g.boxes) {
    bool meet = true;
    for (int a = 0; a < 3; ++a) meet = meet && lo[a] <= b.hi[a] && b.lo[a] <= hi[a];
    if (meet) cls |= b.kind;
  }
  return cls;
}

int plan_step(const PlanGeom &g, StepPlan *out, std::string *err) {
  const int R = g.R, kc = g.kc;
  if (R < 2 || R > 16 || kc < 1 || g.vrows < 2 || g.vrows > R || g.n[0] < 1 || g.n[1] < 1 || g.n[2] < 1 ||
      g.n[0] >= (1 << 20) - 256 || g.n[1] >= (1 << 20)) {
    if (err) *err = "step plan: bad geometry";
    return -1;
  }
  for (int c = 0; c < 8; ++c)
    if (g.rows[c] < 2 || g.rows[c] > R || g.hrows[c] < 2 || g.hrows[c] > R) {
      if (err) *err = "step plan: bad rows per block";
      return -1;
    }
  // cut points per axis: CPML slab edges (a tile is outside the low slab if
  // it starts at L0, outside the high one if its halo stays below n - L1),
  // and the planes just outside each box
  std::vector<int> cuts[3];
  for (int a = 0; a < 3; ++a) {
    if (g.L[a][0] > 0) cuts[a].push_back(g.L[a][0]);
    if (g.L[a][1] > 0) cuts[a].push_back(g.n[a] - g.L[a][1] - 1);
    for (const PlanBox &b : g.boxes) {
      cuts[a].push_back(b.lo[a] - 1);
      cuts[a].push_back(b.hi[a] + 1);
    }
  }
  const int N[3] = {g.n[0] + 1, g.n[1] + 1, g.n[2] + 1};
  const int Wx = kWave - 1, Wh = kWave / 2 - 1;
  // slivers narrower than these fold into a neighbour; wider pieces (a CPML
  // slab with the faces next to it) keep tiles of their own
  std::vector<Seg> sx = segments(cuts[0], N[0], 8);
  std::vector<Seg> sy = segments(cuts[1], N[1], std::max(2, (R - 1) / 2));
  std::vector<Seg> sz = segments(cuts[2], N[2], std::max(2, std::min(kc, 8)));
  // row bands: at most R - 1 rows; with four-cell tiles on smaller blocks, a
  // whole number of their bands, so that no sliver band is left
  int Wy = R - 1;
  if (g.vec && g.vrows < R) Wy = std::max(g.vrows - 1, (R - 1) / (g.vrows - 1) * (g.vrows - 1));
  // the columns of [a, b) four-cell tiles may own: 16-byte aligned bounds,
  // the halo column of the last tile inside every x sweep (x0 + wx + 1 <=
  // nx), and off the x CPML slabs, which a piece holds when a slab narrower
  // than a sliver was folded into it (those columns stay in one-cell tiles)
  auto vec_span = [&](int a, int b, int *a4, int *b4) {
    if (g.L[0][0] > 0) a = std::max(a, g.L[0][0]);
    if (g.L[0][1] > 0) b = std::min(b, g.n[0] - g.L[0][1] - 1);
    *a4 = (a + 3) & ~3;
    *b4 = std::min(b, g.n[0] - 1) & ~3;
    return *b4 - *a4 >= 2 * Wx;
  };
  // a piece wide enough for four-cell tiles moves its bounds onto 16-byte
  // boundaries, its neighbours taking the columns in between
  if (g.vec)
    for (size_t q = 0; q < sx.size(); ++q) {
      int a4, b4;
      if (!vec_span(sx[q].lo, sx[q].hi, &a4, &b4)) continue;
      // (a bound more than 3 columns in is a CPML clip: the piece keeps
      // those columns for its own one-cell tiles)
      if (q > 0 && a4 > sx[q].lo && a4 - sx[q].lo <= 3) { sx[q - 1].hi = a4; sx[q].lo = a4; }
      if (q + 1 < sx.size() && b4 < sx[q].hi && sx[q].hi - b4 <= 3) {
        sx[q + 1].lo = b4;
        sx[q].hi = b4;
      }
    }
  // x tiles: a piece wide enough becomes four-cell tiles between 16-byte
  // aligned bounds; what the rounding leaves runs as 63-column tiles
  std::vector<XTile> xt;
  for (const Seg &s : sx) {
    int a = s.lo, b = s.hi;
    int a4, b4;
    if (g.vec && vec_span(a, b, &a4, &b4)) {
      std::vector<Seg> p;
      if (a4 > a) split(a, a4, Wx, 1, &p);
      for (const Seg &q : p) xt.push_back({q.lo, q.hi, false, q.hi - q.lo <= Wh});
      p.clear();
      split(a4, b4, 4 * Wx, 4, &p);
      for (const Seg &q : p) xt.push_back({q.lo, q.hi, true, false});
      p.clear();
      if (b > b4) split(b4, b, Wx, 1, &p);
      for (const Seg &q : p) xt.push_back({q.lo, q.hi, false, q.hi - q.lo <= Wh});
      continue;
    }
    std::vector<Seg> p;
    split(a, b, Wx, 1, &p);
    for (const Seg &q : p) xt.push_back({q.lo, q.hi, false, q.hi - q.lo <= Wh});
  }
  std::vector<Seg> yt, zt;
  for (const Seg &s : sy) split(s.lo, s.hi, Wy, 1, &yt);
  for (const Seg &s : sz) split(s.lo, s.hi, kc, 1, &zt);
  // tiles in z order of chunks, rows, columns (x fastest); plain tiles merge
  // with the plain tiles above them into runs of up to 64 planes
  for (int c = 0; c < kPlanClasses; ++c) {
    out->tiles[c].clear();
    out->owned[c] = out->psi[c] = out->disp[c][0] = out->disp[c][1] = 0.0;
  }
  // cells of [lo, hi) whose coordinate lies in [a, b)
  auto overlap = [](int lo, int hi, int a, int b) {
    return (double)std::max(0, std::min(hi, b) - std::max(lo, a));
  };
  // psi values and pole-box cells of the owned box [x0, x0+wx) x [y0, y0+wy)
  // x [k0, k1): per component its sweep (E_c: cell along c, node along the
  // others; H_c: node along c, cell along the others), per stretched axis its
  // slab layers (node positions: below L0 or from n - L1 + 1; cell positions:
  // below L0 or from n - L1), as the step kernels select them
  auto state_counts = [&](int cls, int x0, int wx, int y0, int wy, int k0, int k1) {
    const int lo[3] = {x0, y0, k0}, hi[3] = {x0 + wx, y0 + wy, k1};
    for (int comp = 0; comp < 6; ++comp) {
      const int d = comp % 3;
      const bool e = comp < 3;
      double span[3];
      int send[3];
      for (int a = 0; a < 3; ++a) {
        const bool node = e ? a != d : a == d;
        send[a] = node ? g.n[a] + 1 : g.n[a];
        span[a] = overlap(lo[a], hi[a], 0, send[a]);
      }
      for (int a = 0; a < 3; ++a) {
        if (a == d || (g.L[a][0] == 0 && g.L[a][1] == 0)) continue;
        const int hs = e ? g.n[a] - g.L[a][1] + 1 : g.n[a] - g.L[a][1];
        const double in = overlap(lo[a], hi[a], 0, g.L[a][0]) + overlap(lo[a], hi[a], hs, send[a]);
        out->psi[cls] += in * span[(a + 1) % 3] * span[(a + 2) % 3];
      }
      for (const PlanBox &b : g.boxes) {
        if (b.kind != 2 || b.comp != comp) continue;
        double v = 1.0;
        for (int a = 0; a < 3; ++a)
          v *= overlap(lo[a], hi[a], std::max(b.lo[a], 0), std::min(b.hi[a] + 1, send[a]));
        out->disp[cls][e ? 0 : 1] += v;
      }
    }
  };
  struct Pending { int cls; int x0, wx, y0, wy, k0, k1; bool open; };
  const size_t nxy = xt.size() * yt.size() * 4;   // up to 4 scalar pieces per x tile
  std::vector<Pending> run(nxy, Pending{-1, 0, 0, 0, 0, 0, 0, false});
  auto owned_cells = [&](int x0, int wx, int y0, int wy, int k0, int k1) {
    const double ox = std::max(0, std::min(x0 + wx, g.n[0]) - x0);
    const double oy = std::max(0, std::min(y0 + wy, g.n[1]) - y0);
    const double oz = std::max(0, std::min(k1, g.n[2]) - k0);
    return ox * oy * oz;
  };
  auto emit = [&](int cls, int x0, int wx, int y0, int wy, int k0, int k1) {
    out->owned[cls] += owned_cells(x0, wx, y0, wy, k0, k1);
    state_counts(cls, x0, wx, y0, wy, k0, k1);
    const int E = cls < 8 ? g.rows[cls] : g.vrows;
    if (E >= R) {
      out->tiles[cls].push_back(step_tile(x0, wx, y0, wy, k0, k1));
      return;
    }
    // a class on smaller blocks: the tile's rows in near-equal bands of at
    // most E - 1
    std::vector<Seg> bands;
    split(y0, y0 + wy, E - 1, 1, &bands);
    for (const Seg &b : bands)
      out->tiles[cls].push_back(step_tile(x0, wx, b.lo, b.hi - b.lo, k0, k1));
  };
  auto flush = [&](Pending &p) {
    if (p.open) emit(p.cls, p.x0, p.wx, p.y0, p.wy, p.k0, p.k1);
    p.open = false;
  };
  // half-width tiles: per z tile and y piece, the class of the x piece's
  // columns over the whole y piece (halo included), then bands of at most
  // 2E - 1 rows for the class's E waves per block
  for (const Seg &z : zt)
    for (const Seg &y : sy)
      for (const XTile &x : xt) {
        if (!x.half) continue;
        const int lo[3] = {x.lo, y.lo, z.lo}, hi[3] = {x.hi, y.hi, z.hi};
        const int c = plan_classify(g, lo, hi);
        const int cls = kPlanHalf + c;
        std::vector<Seg> bands;
        split(y.lo, y.hi, 2 * g.hrows[c] - 1, 1, &bands);
        for (const Seg &b : bands) {
          const int wy = b.hi - b.lo;
          out->owned[cls] += owned_cells(x.lo, x.hi - x.lo, b.lo, wy, z.lo, z.hi);
          state_counts(cls, x.lo, x.hi - x.lo, b.lo, wy, z.lo, z.hi);
          out->tiles[cls].push_back(step_tile(x.lo, x.hi - x.lo, b.lo, wy, z.lo, z.hi));
        }
      }
  for (const Seg &z : zt) {
    for (size_t by = 0; by < yt.size(); ++by) {
      const Seg &y = yt[by];
      for (size_t bx = 0; bx < xt.size(); ++bx) {
        const XTile &x = xt[bx];
        if (x.half) continue;
        // the pieces of this x tile: itself, or its 63-column split when a
        // four-cell tile is not plain
        Seg pieces[4];
        int np = 0, pc[4];
        const int lo[3] = {x.lo, y.lo, z.lo}, hi[3] = {x.hi, y.hi, z.hi};
        const int c0 = plan_classify(g, lo, hi);
        if (x.vec && c0 == 0) {
          pieces[0] = {x.lo, x.hi};
          pc[0] = 8;
          np = 1;
        } else if (x.vec) {
          std::vector<Seg> p;
          split(x.lo, x.hi, Wx, 1, &p);
          for (const Seg &q : p) {
            const int l2[3] = {q.lo, y.lo, z.lo}, h2[3] = {q.hi, y.hi, z.hi};
            pieces[np] = q;
            pc[np++] = plan_classify(g, l2, h2);
          }
        } else {
          pieces[0] = {x.lo, x.hi};
          pc[0] = c0;
          np = 1;
        }
        for (int q = 0; q < 4; ++q) {
          Pending &p = run[(by * xt.size() + bx) * 4 + q];
          if (q >= np) { flush(p); continue; }
          const int cls = pc[q];
          const Seg &s = pieces[q];
          const bool mergeable = cls == 0 || cls == 8;
          if (p.open && mergeable && p.cls == cls && p.x0 == s.lo && p.wx == s.hi - s.lo &&
              p.k1 == z.lo && z.hi - p.k0 <= 64) {
            p.k1 = z.hi;
            continue;
          }
          flush(p);
          if (mergeable) {
            p = Pending{cls, s.lo, s.hi - s.lo, y.lo, y.hi - y.lo, z.lo, z.hi, true};
          } else {
            emit(cls, s.lo, s.hi - s.lo, y.lo, y.hi - y.lo, z.lo, z.hi);
          }
        }
      }
    }
  }
  for (Pending &p : run) flush(p);
  // the merged runs were emitted when they closed: order each class's tiles
  // by (k0, y0, x0) so that neighbouring tiles sit next to each other
  for (int c = 0; c < kPlanClasses; ++c)
    std::stable_sort(out->tiles[c].begin(), out->tiles[c].end(),
                     [](const StepTile &a, const StepTile &b) {
                       if (a.k0 != b.k0) return a.k0 < b.k0;
                       const int ya = a.y & 0xfffff, yb = b.y & 0xfffff;
                       if (ya != yb) return ya < yb;
                       return (a.x & 0xfffff) < (b.x & 0xfffff);
                     });
  return 0;
}

}  // namespace sfd_host

// The CPU tests' view of the plan: geom = {nx, ny, nz, Lx0, Lx1, Ly0, Ly1,
// Lz0, Lz1, R, kc, vec, vrows}; rows[16] rows per block of classes 0-7, then
// waves per block of their half-width variants (classes 9-16); boxes
// [nbox][7] = lo xyz, hi xyz (inclusive), kind (2 poles, 4 sources).  Writes
// up to cap tiles as [cls, x0, wx, y0, wy, k0, k1] and the count to *ntiles
// (all of them, also when more than cap).
extern "C" int sfd_plan_step(const int32_t *geom, const int32_t *rows, int32_t nbox,
                             const int32_t *boxes, int32_t *tiles, int64_t cap,
                             int64_t *ntiles) {
  using namespace sfd_host;
  if (!geom || !rows || !ntiles || (nbox > 0 && !boxes)) return -1;
  PlanGeom g;
  for (int a = 0; a < 3; ++a) {
    g.n[a] = geom[a];
    g.L[a][0] = geom[3 + 2 * a];
    g.L[a][1] = geom[4 + 2 * a];
  }
  g.R = geom[9];
  g.kc = geom[10];
  g.vec = geom[11] != 0;
  g.vrows = geom[12];
  for (int c = 0; c < 8; ++c) {
    g.rows[c] = rows[c];
    g.hrows[c] = rows[8 + c];
  }
  for (int b = 0; b < nbox; ++b) {
    PlanBox pb;
    for (int a = 0; a < 3; ++a) {
      pb.lo[a] = boxes[b * 7 + a];
      pb.hi[a] = boxes[b * 7 + 3 + a];
    }
    pb.kind = boxes[b * 7 + 6];
    g.boxes.push_back(pb);
  }
  StepPlan p;
  if (plan_step(g, &p, nullptr)) return -1;
  int64_t n = 0;
  for (int c = 0; c < kPlanClasses; ++c)
    for (const sfd::StepTile &t : p.tiles[c]) {
      if (n < cap && tiles) {
        int32_t *o = tiles + n * 7;
        o[0] = c;
        o[1] = t.x & 0xfffff; o[2] = t.x >> 20;
        o[3] = t.y & 0xfffff; o[4] = t.y >> 20;
        o[5] = t.k0; o[6] = t.k1;
      }
      ++n;
    }
  *ntiles = n;
  return 0;
}
