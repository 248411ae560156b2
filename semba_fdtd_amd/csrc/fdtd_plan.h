// fdtd_plan.h -- the tile plan of the one-launch step (fdtd_plan.cpp).
#pragma once
#include <string>
#include <vector>

#include "fdtd_internal.h"

namespace sfd_host {

// tile classes: bits 0 CPML, 1 poles, 2 sources (0..7, one k_step variant
// each); 8: plain tiles run four x cells per lane (k_step4); 9..16: the
// half-width variant of class c - 9 (k_step with 32 lanes per row, for x
// pieces of at most 31 columns)
constexpr int kPlanClasses = 17;
constexpr int kPlanVec = 8;
constexpr int kPlanHalf = 9;

// a box of cells [lo, hi] (inclusive) that tiles meeting it must handle:
// kind 2 a pole-state box (of field component comp, 0-5), 4 a TF/SF face or
// the nodal cells of a half (comp -1)
struct PlanBox {
  int lo[3], hi[3];
  int kind;
  int comp = -1;
};

struct PlanGeom {
  int n[3];          // nx, ny, nz (the thread domain is [0, n] per axis)
  int L[3][2];       // CPML layers per side
  int R;             // rows per block of the step (tiles own up to R - 1 rows)
  int kc;            // planes per z chunk (tiles with sources own at most kc)
  int rows[8];       // rows per block of each class (<= R)
  int hrows[8];      // waves per block of each half-width class (<= R; a tile
                     // owns up to 2 hrows - 1 rows)
  int vrows;         // rows per block of the four-cell kernel (<= R)
  bool vec;          // plain tiles may run four x cells per lane
  std::vector<PlanBox> boxes;
};

struct StepPlan {
  std::vector<sfd::StepTile> tiles[kPlanClasses];
  double owned[kPlanClasses];   // Yee cells (i < nx, j < ny, k < nz) owned
  double psi[kPlanClasses];     // CPML psi values the class's tiles update
  double disp[kPlanClasses][2]; // pole-state box cells they update (E, H side)
};

// class bits of a tile whose cells, halo included, span [lo, hi]
int plan_classify(const PlanGeom &g, const int lo[3], const int hi[3]);
int plan_step(const PlanGeom &g, StepPlan *out, std::string *err);

}  // namespace sfd_host
