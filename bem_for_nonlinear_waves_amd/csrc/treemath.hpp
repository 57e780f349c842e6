// Octree cell arithmetic and the multipole acceptance criterion, shared by the host plan builder (plan.cpp)
// and the device tree builder (tree.hip), so that both put every face in the same box and accept the same
// box pairs. The operations are those of the reference's Buckets (lib/include/lib_spatial_partitioning.hpp:
// 223-330: cell index floor((x - xmin) / dL) clamped to the grid, InsideQ without the clamp, children of
// side dL (0.5 + 1e-13) on the parent's cells) and of isFar (lib/include/moments.hpp:948-954), with
// std::hypot's three-argument form written out as libstdc++ evaluates it. No fused multiply-adds: tree.hip
// is compiled with -ffp-contract=off and the host objects for an x86-64 level without FMA.
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#define BEMFMM_TREE_HD __host__ __device__
#else
#define BEMFMM_TREE_HD
#endif

namespace bemfmm {

struct BoxGeom {
  double bmin[3];
  double X[3];    // centre: mean of the bounds
  double dL;      // cell side
  double inv_dL;  // 1 / dL
  int size[3];    // cells per axis
};

// a box over [lo, hi] with cells of side dl
BEMFMM_TREE_HD inline void box_init(const double lo[3], const double hi[3], double dl, BoxGeom& g) {
  for (int a = 0; a < 3; ++a) {
    g.bmin[a] = lo[a];
    g.X[a] = (lo[a] + hi[a]) / 2.;
    double s = ceil((hi[a] - lo[a]) / dl);
    if (s <= 0.) s = 1.;
    g.size[a] = (int)s;
  }
  g.dL = dl;
  g.inv_dL = 1.0 / dl;
}

BEMFMM_TREE_HD inline int box_cells(const BoxGeom& g) { return g.size[0] * g.size[1] * g.size[2]; }

// cell of point x, each index clamped to the grid
BEMFMM_TREE_HD inline int box_cell_of(const BoxGeom& g, const double x[3]) {
  int ijk[3];
  for (int a = 0; a < 3; ++a) {
    int i = (int)floor((x[a] - g.bmin[a]) * g.inv_dL);
    ijk[a] = i < 0 ? 0 : (i > g.size[a] - 1 ? g.size[a] - 1 : i);
  }
  return (ijk[0] * g.size[1] + ijk[1]) * g.size[2] + ijk[2];
}

BEMFMM_TREE_HD inline bool box_inside(const BoxGeom& g, const double x[3]) {
  for (int a = 0; a < 3; ++a) {
    const int i = (int)floor((x[a] - g.bmin[a]) * g.inv_dL);
    if (i < 0 || i >= g.size[a]) return false;
  }
  return true;
}

// the child box on cell `cell` of p
BEMFMM_TREE_HD inline void box_child(const BoxGeom& p, int cell, BoxGeom& c) {
  const int k = cell % p.size[2], j = (cell / p.size[2]) % p.size[1], i = cell / (p.size[2] * p.size[1]);
  const int ijk[3] = {i, j, k};
  double lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = p.dL * ijk[a] + p.bmin[a];
    hi[a] = p.dL * (ijk[a] + 1) + p.bmin[a];
  }
  box_init(lo, hi, p.dL * (0.5 + 1e-13), c);
}

// sqrt(x^2 + y^2 + z^2) scaled by the largest magnitude, in libstdc++'s order of operations
BEMFMM_TREE_HD inline double hypot3(double x, double y, double z) {
  x = fabs(x);
  y = fabs(y);
  z = fabs(z);
  const double a = x < y ? (y < z ? z : y) : (x < z ? z : x);
  if (a == 0.) return 0.;
  const double u = x / a, v = y / a, w = z / a;
  return a * sqrt(u * u + v * v + w * w);
}

// isFar: boxes of cell sides dL and internal offsets off, centres Xa and Xb
BEMFMM_TREE_HD inline bool boxes_far(const double Xa[3], double dLa, double offa, const double Xb[3], double dLb,
                                     double offb, double theta) {
  const double ea = dLa + 2.0 * offa, eb = dLb + 2.0 * offb;
  const double eff = ea < eb ? eb : ea;
  const double d = hypot3(Xb[0] - Xa[0], Xb[1] - Xa[1], Xb[2] - Xa[2]);
  return theta > sqrt(3.0) * eff / d;
}

}  // namespace bemfmm
