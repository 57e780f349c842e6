// Host-side plan builder for one FMM-GMRES BVP (see plan.hpp).
//
// Follows the reference setup path (bem/BEM_solveBVP_GMRES_FMM.hpp:3056-3450 and
// lib/include/moments.hpp:1239-1541): coordinate scaling, per-face P2M sources,
// the Buckets octree and its MAC/interaction lists, per-target near lists and
// the translation pairs. Arithmetic that decides tree membership and the MAC is
// kept bit-identical to the reference's (cell index = floor((x-xmin)*inv_dL),
// child cell size = dL*(0.5+1e-13), centre = mean of bounds, Norm = hypot).
#include <omp.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>

#include "nearquad.hpp"
#include "plan.hpp"

namespace bemfmm {

// Block GEMV rows are padded to a multiple of 4 columns, so that every row starts 16-byte aligned
// and the GEMV streams them with 16-byte vector loads.
static inline int pad_cols(int n) { return (n + 3) & ~3; }
namespace {

using V3 = std::array<double, 3>;
inline double hyp3(double a, double b, double c) { return hypot3(a, b, c); }

double now() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

// Dunavant rules for the P2M sources (lib/include/dunavant_rules.hpp:18-46), {b0,b1,b2,w}, sum w = 1/2.
const std::vector<std::array<double, 4>>& dunavant(int q) {
  static const std::vector<std::array<double, 4>> d1 = {{1. / 3., 1. / 3., 1. / 3., 0.5}};
  static const std::vector<std::array<double, 4>> d3 = {
      {2. / 3., 1. / 6., 1. / 6., 1. / 6.}, {1. / 6., 2. / 3., 1. / 6., 1. / 6.}, {1. / 6., 1. / 6., 2. / 3., 1. / 6.}};
  static const std::vector<std::array<double, 4>> d6 = {
      {0.816847572980459, 0.091576213509771, 0.091576213509771, 0.054975871827661},
      {0.091576213509771, 0.816847572980459, 0.091576213509771, 0.054975871827661},
      {0.091576213509771, 0.091576213509771, 0.816847572980459, 0.054975871827661},
      {0.108103018168070, 0.445948490915965, 0.445948490915965, 0.111690794839006},
      {0.445948490915965, 0.108103018168070, 0.445948490915965, 0.111690794839006},
      {0.445948490915965, 0.445948490915965, 0.108103018168070, 0.111690794839006}};
  static const std::vector<std::array<double, 4>> d7 = {
      {1. / 3., 1. / 3., 1. / 3., 0.1125},
      {0.059715871789770, 0.470142064105115, 0.470142064105115, 0.066197076394253},
      {0.470142064105115, 0.059715871789770, 0.470142064105115, 0.066197076394253},
      {0.470142064105115, 0.470142064105115, 0.059715871789770, 0.066197076394253},
      {0.797426985353087, 0.101286507323456, 0.101286507323456, 0.062969590272414},
      {0.101286507323456, 0.797426985353087, 0.101286507323456, 0.062969590272414},
      {0.101286507323456, 0.101286507323456, 0.797426985353087, 0.062969590272414}};
  switch (q) {
  case 1: return d1;
  case 3: return d3;
  case 6: return d6;
  case 7: return d7;
  default: throw std::runtime_error("p2m_quadrature_points must be one of 1, 3, 6, 7");
  }
}

struct TBox {
  int level = 0, parent = -1;
  BoxGeom g;
  std::vector<int> cells;
  std::vector<std::vector<int>> cell_data;
  std::vector<int> data;
  double off = 0.;
  bool has_children = false;
  int id = -1;  // final level-major id
  std::vector<int> near, m2l;
  void init(const BoxGeom& geom) {
    g = geom;
    cells.assign(box_cells(g), -1);
    cell_data.assign(box_cells(g), {});
  }
};

struct Builder {
  Plan& pl;
  std::vector<TBox> B;
  explicit Builder(Plan& p) : pl(p) {}
  const double* pole(int f) const { return &pl.pole[3 * (size_t)f]; }

  void add(int b, int f) {
    B[b].data.push_back(f);
    B[b].cell_data[box_cell_of(B[b].g, pole(f))].push_back(f);
  }
  void grow(int b) {
    // generateTree with the grow condition of BEM_solveBVP_GMRES_FMM.hpp:1199
    const bool cond = !B[b].data.empty() && (int)B[b].data.size() > pl.opt.leaf_max && B[b].level < pl.opt.max_level;
    if (!cond) return;
    const int ncell = (int)B[b].cells.size();
    for (int c = 0; c < ncell; ++c) {
      if (B[b].cell_data[c].empty()) continue;
      BoxGeom cg;
      box_child(B[b].g, c, cg);
      TBox ch;
      ch.init(cg);
      ch.parent = b;
      ch.level = B[b].level + 1;
      const int cid = (int)B.size();
      B.push_back(std::move(ch));
      for (int f : B[b].cell_data[c])
        if (box_inside(B[cid].g, pole(f))) add(cid, f);
      B[b].cells[c] = cid;
      B[b].has_children = true;
      grow(cid);
    }
  }
  bool is_far(int a, int b) const {
    return boxes_far(B[a].g.X, B[a].g.dL, B[a].off, B[b].g.X, B[b].g.dL, B[b].off, pl.opt.theta);
  }
  int deepest(const double* x) const {
    int cur = 0;
    while (B[cur].has_children) {
      const int nx = B[cur].cells[box_cell_of(B[cur].g, x)];
      if (nx < 0) return cur;
      cur = nx;
    }
    return cur;
  }
};

// Solid harmonics in double: NR_n^m (m >= 0, all n <= P) by the recurrences
// NR_m^m = (x+iy)/(2m) NR_{m-1}^{m-1}, NR_n^m = ((2n-1) z NR_{n-1}^m - r^2 NR_{n-2}^m)/((n+m)(n-m)).
struct Harm {
  int P;
  std::vector<double> re, im;  // idx(n,m) = n*(n+1)/2 + m
  std::vector<double> cnm;     // (-1)^m sqrt((n-m)!(n+m)!)
  explicit Harm(int P_) : P(P_), re((P_ + 2) * (P_ + 3) / 2), im((P_ + 2) * (P_ + 3) / 2), cnm((P_ + 2) * (P_ + 3) / 2) {
    for (int n = 0; n <= P + 1; ++n)
      for (int m = 0; m <= n; ++m) {
        double f1 = 1, f2 = 1;
        for (int k = 2; k <= n - m; ++k) f1 *= k;
        for (int k = 2; k <= n + m; ++k) f2 *= k;
        const double v = std::sqrt(f1 * f2);
        cnm[t(n, m)] = (m & 1) ? -v : v;
      }
  }
  static int t(int n, int m) { return n * (n + 1) / 2 + m; }
  void eval(double x, double y, double z, int L) {
    const double r2 = x * x + y * y + z * z;
    re[0] = 1.0;
    im[0] = 0.0;
    for (int m = 1; m <= L; ++m) {
      const double pr = re[t(m - 1, m - 1)], pi = im[t(m - 1, m - 1)];
      const double s = 1.0 / (2.0 * m);
      re[t(m, m)] = s * (x * pr - y * pi);
      im[t(m, m)] = s * (x * pi + y * pr);
    }
    for (int m = 0; m <= L; ++m)
      for (int n = m + 1; n <= L; ++n) {
        double ar = (2 * n - 1) * z * re[t(n - 1, m)], ai = (2 * n - 1) * z * im[t(n - 1, m)];
        if (n - 2 >= m) {
          ar -= r2 * re[t(n - 2, m)];
          ai -= r2 * im[t(n - 2, m)];
        }
        const double s = 1.0 / ((double)(n + m) * (n - m));
        re[t(n, m)] = s * ar;
        im[t(n, m)] = s * ai;
      }
  }
  // NR_n^q for any q (NR_n^{-m} = (-1)^m conj(NR_n^m)); zero if |q| > n or n < 0
  void nr(int n, int q, double& r, double& i) const {
    if (n < 0 || q > n || -q > n) {
      r = i = 0;
      return;
    }
    if (q >= 0) {
      r = re[t(n, q)];
      i = im[t(n, q)];
    } else {
      const int m = -q;
      const double s = (m & 1) ? -1.0 : 1.0;
      r = s * re[t(n, m)];
      i = -s * im[t(n, m)];
    }
  }
};

}  // namespace

void plan_geometry(Plan& pl, const Options& opt, int n_nodes, const double* X, int n_faces, const int32_t* tri,
                   const int32_t* face_dof, int n_dof, const int32_t* dof_kind, const int32_t* dof_point) {
  if (opt.p < 1 || opt.p > 20) throw std::runtime_error("expansion order p must be in [1, 20]");
  if (n_nodes <= 0 || n_faces <= 0 || n_dof <= 0) throw std::runtime_error("empty geometry");
  pl.opt = opt;
  pl.P = opt.p;
  pl.C = (opt.p + 1) * (opt.p + 1);
  pl.n_nodes = n_nodes;
  pl.n_faces = n_faces;
  pl.n_dof = n_dof;
  if (opt.npf != 3 && opt.npf != 6) throw std::runtime_error("nodes per face must be 3 (linear) or 6 (quadratic)");
  const int npf = pl.npf = opt.npf;
  for (int64_t e = 0; e < (int64_t)npf * n_faces; ++e) {
    if (tri[e] < 0 || tri[e] >= n_nodes) throw std::runtime_error("triangle node index out of range");
    const bool midpoint = npf == 6 && e % 6 >= 3;  // an edge midpoint may carry no DOF (-1)
    if ((face_dof[e] < 0 && !(midpoint && face_dof[e] == -1)) || face_dof[e] >= n_dof)
      throw std::runtime_error("face DOF index out of range");
  }
  for (int i = 0; i < n_dof; ++i) {
    if (dof_kind[i] < 0 || dof_kind[i] > 3)
      throw std::runtime_error("dof_kind must be 0 (Dirichlet), 1 (Neumann), 2 (CORNER Neumann) or 3 (ROBIN)");
    if (dof_point[i] < 0 || dof_point[i] >= n_nodes) throw std::runtime_error("dof_point out of range");
  }
  pl.dof_kind.assign(dof_kind, dof_kind + n_dof);
  pl.dof_point.assign(dof_point, dof_point + n_dof);
  const double t0 = now();

  // --- coordinate scaling (Network::computeCharacteristicLength / applyScaling) ---
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int i = 0; i < n_nodes; ++i)
    for (int a = 0; a < 3; ++a) {
      const double v = X[3 * i + a];
      if (!std::isfinite(v)) throw std::runtime_error("non-finite node coordinate");
      lo[a] = std::min(lo[a], v);
      hi[a] = std::max(hi[a], v);
    }
  pl.Lscale = std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) + (hi[2] - lo[2]) * (hi[2] - lo[2]));
  if (!(pl.Lscale > 1e-10)) throw std::runtime_error("degenerate geometry (zero bounding box)");
  pl.Xs.resize(3 * (size_t)n_nodes);
  double slo[3] = {1e300, 1e300, 1e300}, shi[3] = {-1e300, -1e300, -1e300};
  for (int i = 0; i < n_nodes; ++i)
    for (int a = 0; a < 3; ++a) {
      const double v = X[3 * i + a] / pl.Lscale;
      pl.Xs[3 * i + a] = v;
      slo[a] = std::min(slo[a], v);
      shi[a] = std::max(shi[a], v);
    }
  pl.scale = hyp3(shi[0] - slo[0], shi[1] - slo[1], shi[2] - slo[2]);
  pl.near_region = pl.scale / 25.;

  // --- faces and P2M sources (GMRES_FMM.hpp:1799-1937) ---
  const auto& rule = dunavant(opt.q);
  pl.fX.resize(3 * (size_t)npf * n_faces);
  pl.fcent.resize(3 * (size_t)n_faces);
  pl.fcr.resize(3 * (size_t)n_faces);
  pl.fJ.resize(n_faces);
  pl.fv.assign(tri, tri + (size_t)npf * n_faces);
  pl.fdof.assign(face_dof, face_dof + (size_t)npf * n_faces);
  pl.pole.resize(3 * (size_t)n_faces);
  pl.maxoff.resize(n_faces);
  int degenerate = 0;
#pragma omp parallel for schedule(static) reduction(| : degenerate)
  for (int f = 0; f < n_faces; ++f) {
    V3 x[3];
    for (int k = 0; k < npf; ++k)
      for (int a = 0; a < 3; ++a) {
        const double v = pl.Xs[3 * tri[(int64_t)npf * f + k] + a];
        if (k < 3) x[k][a] = v;
        pl.fX[(size_t)3 * npf * f + 3 * k + a] = v;
      }
    const V3 e1 = {x[1][0] - x[0][0], x[1][1] - x[0][1], x[1][2] - x[0][2]};
    const V3 e2 = {x[2][0] - x[0][0], x[2][1] - x[0][1], x[2][2] - x[0][2]};
    const V3 cr = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const double J = hyp3(cr[0], cr[1], cr[2]);
    if (!(J > 0)) degenerate = 1;
    pl.fJ[f] = J;
    double* pf = &pl.pole[3 * (size_t)f];
    for (int a = 0; a < 3; ++a) {
      pl.fcr[3 * f + a] = J * (cr[a] / J);
      pl.fcent[3 * f + a] = (x[0][a] + x[1][a] + x[2][a]) / 3.;
      // the tree position: Dot({1/3,1/3,1/3}, X012) for linear faces, the centroid (a+b+c)/3 for quadratic ones
      pf[a] = npf == 3 ? (1. / 3.) * x[0][a] + (1. / 3.) * x[1][a] + (1. / 3.) * x[2][a] : pl.fcent[3 * f + a];
    }
    double mo = 0.;
    for (const auto& r : rule) {
      V3 xq;
      if (npf == 3) {
        for (int a = 0; a < 3; ++a) xq[a] = r[0] * x[0][a] + r[1] * x[1][a] + r[2] * x[2][a];
      } else {
        double cr6[3];
        tri6_geom(&pl.fX[(size_t)18 * f], r[0], r[1], xq.data(), cr6);
      }
      mo = std::max(mo, hyp3(xq[0] - pf[0], xq[1] - pf[1], xq[2] - pf[2]));
    }
    pl.maxoff[f] = mo;
  }
  if (degenerate) throw std::runtime_error("degenerate triangle (zero area)");

  // --- root box (initializeBucket GMRES_FMM.hpp:1178-1201) ---
  {
    double blo[3], bhi[3];
    const double dL = 1.05 * pl.scale / 2.;
    for (int a = 0; a < 3; ++a) {
      const double c = (slo[a] + shi[a]) / 2.;
      blo[a] = c - dL;
      bhi[a] = c + dL;
    }
    box_init(blo, bhi, dL, pl.root);
    pl.root_side = bhi[0] - blo[0];
  }
  for (int f = 0; f < n_faces; ++f)
    if (!box_inside(pl.root, &pl.pole[3 * (size_t)f])) throw std::runtime_error("face outside the root box");
  pl.t_tree = now() - t0;
}

void plan_tree_lists(Plan& pl) {
  const Options& opt = pl.opt;
  const int n_nodes = pl.n_nodes, n_faces = pl.n_faces;
  const double t0 = now();
  // --- octree (generateTree) ---
  Builder bd(pl);
  {
    TBox root;
    root.init(pl.root);
    bd.B.push_back(std::move(root));
  }
  for (int f = 0; f < n_faces; ++f) bd.add(0, f);
  bd.grow(0);
  auto& Bx = bd.B;
  // level lists in preorder (rebuildHierarchyLists)
  std::vector<std::vector<int>> lv(opt.max_level + 1);
  {
    // preorder with children in cell order
    std::function<void(int)> visit = [&](int b) {
      if (Bx[b].data.empty()) throw std::runtime_error("empty octree box (face on a box boundary)");
      lv[Bx[b].level].push_back(b);
      for (int c : Bx[b].cells)
        if (c >= 0) visit(c);
    };
    visit(0);
  }
  pl.n_levels = 0;
  for (int l = 0; l <= opt.max_level; ++l)
    if (!lv[l].empty()) pl.n_levels = l + 1;
  pl.level_ptr.assign(pl.n_levels + 1, 0);
  int nb = 0;
  for (int l = 0; l < pl.n_levels; ++l) {
    for (int b : lv[l]) Bx[b].id = nb++;
    pl.level_ptr[l + 1] = nb;
  }
  pl.n_boxes = nb;
  std::vector<int> id2b(nb);
  for (int b = 0; b < (int)Bx.size(); ++b) id2b[Bx[b].id] = b;
  pl.box_level.resize(nb);
  pl.box_parent.resize(nb);
  pl.box_leaf.resize(nb);
  pl.box_center.resize(3 * (size_t)nb);
  for (int i = 0; i < nb; ++i) {
    const TBox& t = Bx[id2b[i]];
    pl.box_level[i] = t.level;
    pl.box_parent[i] = t.parent >= 0 ? Bx[t.parent].id : -1;
    pl.box_leaf[i] = t.has_children ? 0 : 1;
    for (int a = 0; a < 3; ++a) pl.box_center[3 * i + a] = t.g.X[a];
  }
  // DFS face order; every box's faces are a contiguous range
  pl.face_perm.clear();
  pl.face_perm.reserve(n_faces);
  pl.box_range.assign(2 * (size_t)nb, 0);
  {
    std::function<void(int)> emit = [&](int b) {
      const int64_t s = (int64_t)pl.face_perm.size();
      if (!Bx[b].has_children) {
        pl.face_perm.insert(pl.face_perm.end(), Bx[b].data.begin(), Bx[b].data.end());
      } else {
        // faces of this box that no child kept (reference: lost by the child's InsideQ)
        std::vector<int> inc;
        for (int c : Bx[b].cells)
          if (c >= 0) inc.insert(inc.end(), Bx[c].data.begin(), Bx[c].data.end());
        std::sort(inc.begin(), inc.end());
        for (int f : Bx[b].data)
          if (!std::binary_search(inc.begin(), inc.end(), f)) pl.face_perm.push_back(f);
        for (int c : Bx[b].cells)
          if (c >= 0) emit(c);
      }
      pl.box_range[2 * Bx[b].id] = s;
      pl.box_range[2 * Bx[b].id + 1] = (int64_t)pl.face_perm.size();
    };
    emit(0);
  }
  pl.n_leaves = 0;
  for (int i = 0; i < nb; ++i) pl.n_leaves += pl.box_leaf[i];
  // --- targets: deepest box per node (getBucketAtDeepest) ---
  pl.target_box.resize(n_nodes);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n_nodes; ++i) pl.target_box[i] = Bx[bd.deepest(&pl.Xs[3 * (size_t)i])].id;
  pl.t_tree += now() - t0;

  // --- max_internal_offset (moments.hpp:1323-1337) ---
  const double t1 = now();
  for (int i = nb - 1; i >= 0; --i) {
    TBox& t = Bx[id2b[i]];
    t.off = 0.;
    if (!t.has_children) {
      for (int f : t.data) t.off = std::max(t.off, pl.maxoff[f]);
    } else {
      for (int c : t.cells)
        if (c >= 0) t.off = std::max(t.off, Bx[c].off);
    }
  }
  // --- near / M2L lists (buildNearAndM2LLists, moments.hpp:968-979), level-synchronous ---
  {
    std::vector<std::pair<int, int>> pairs;  // near pairs at the current level (A, B) as raw indices
    for (int a : Bx[0].cells)
      if (a >= 0)
        for (int b : Bx[0].cells)
          if (b >= 0) pairs.push_back({a, b});
    // the root's children: buildNearAndM2LLists(A,B) is entered only when isNear(A,B)
    while (!pairs.empty()) {
      std::vector<std::pair<int, int>> next;
      const int np = (int)pairs.size();
      const int nt = omp_get_max_threads();
      std::vector<std::vector<std::pair<int, int>>> next_t(nt), m2l_t(nt), near_t(nt);
#pragma omp parallel
      {
        const int tid = omp_get_thread_num();
#pragma omp for schedule(dynamic, 64)
        for (int i = 0; i < np; ++i) {
          const int A = pairs[i].first, Bb = pairs[i].second;
          if (bd.is_far(A, Bb)) continue;  // isNear(A,B) is false
          if (A != Bb) near_t[tid].push_back({A, Bb});
          for (int ac : Bx[A].cells) {
            if (ac < 0) continue;
            for (int bc : Bx[Bb].cells) {
              if (bc < 0) continue;
              if (ac != bc && bd.is_far(ac, bc))
                m2l_t[tid].push_back({ac, bc});
              else
                next_t[tid].push_back({ac, bc});
            }
          }
        }
      }
      for (int t = 0; t < nt; ++t) {
        for (auto& pr : near_t[t]) Bx[pr.first].near.push_back(pr.second);
        for (auto& pr : m2l_t[t]) Bx[pr.first].m2l.push_back(pr.second);
        next.insert(next.end(), next_t[t].begin(), next_t[t].end());
      }
      pairs.swap(next);
    }
  }
  // as CSR by box id, ascending
  auto to_csr = [&](std::vector<int> TBox::*list, std::vector<int64_t>& ptr, std::vector<int32_t>& idx) {
    ptr.assign(nb + 1, 0);
    for (int i = 0; i < nb; ++i) ptr[i + 1] = ptr[i] + (int64_t)(Bx[id2b[i]].*list).size();
    idx.assign(ptr[nb], 0);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < nb; ++i) {
      int32_t* o = &idx[ptr[i]];
      int k = 0;
      for (int r : Bx[id2b[i]].*list) o[k++] = Bx[r].id;
      std::sort(o, o + k);
    }
  };
  to_csr(&TBox::near, pl.near_ptr, pl.near_idx);
  to_csr(&TBox::m2l, pl.m2l_ptr, pl.m2l_idx);
  pl.t_lists = now() - t1;
}

void build_plan(Plan& pl, const Options& opt, int n_nodes, const double* X, int n_faces, const int32_t* tri,
                const int32_t* face_dof, int n_dof, const int32_t* dof_kind, const int32_t* dof_point) {
  plan_geometry(pl, opt, n_nodes, X, n_faces, tri, face_dof, n_dof, dof_kind, dof_point);
  plan_tree_lists(pl);
  plan_finish(pl);
}

void plan_level_sides(Plan& pl) {
  pl.level_side.resize(pl.n_levels);
  for (int l = 0; l < pl.n_levels; ++l) pl.level_side[l] = pl.root_side / std::ldexp(1.0, l);
}

void plan_finish(Plan& pl, bool near_on_host) {
  const Options& opt = pl.opt;
  const int npf = pl.npf;
  const int P = pl.P, C = pl.C, n_nodes = pl.n_nodes, n_dof = pl.n_dof, nb = pl.n_boxes;
  const int32_t* face_dof = pl.fdof.data();
  const auto& rule = dunavant(opt.q);
  if ((int)pl.level_side.size() != pl.n_levels) plan_level_sides(pl);
  const double t2 = now();
  // --- near blocks: one per target box (split into <= kMaxRows rows) ---
  if (near_on_host) {
    std::vector<std::vector<int>> box_targets(nb);
    for (int i = 0; i < n_nodes; ++i) box_targets[pl.target_box[i]].push_back(i);
    constexpr int kMaxRows = 32;
    std::vector<int> tboxes;
    for (int b = 0; b < nb; ++b)
      if (!box_targets[b].empty()) tboxes.push_back(b);
    const int ntb = (int)tboxes.size();
    std::vector<std::vector<int64_t>> ranges(ntb);
    std::vector<std::vector<int32_t>> cols(ntb);
    int overlap = 0;  // set inside the parallel loop; reported after it (exceptions must not leave it)
#pragma omp parallel reduction(| : overlap)
    {
      std::vector<int> mark(n_dof, -1);
#pragma omp for schedule(dynamic, 8)
      for (int ti = 0; ti < ntb; ++ti) {
        const int b = tboxes[ti];
        std::vector<std::pair<int64_t, int64_t>> rg;
        auto addbox = [&](int id) { rg.push_back({pl.box_range[2 * id], pl.box_range[2 * id + 1]}); };
        addbox(b);
        for (int64_t k = pl.near_ptr[b]; k < pl.near_ptr[b + 1]; ++k) addbox(pl.near_idx[k]);
        for (int a = pl.box_parent[b]; a >= 0; a = pl.box_parent[a])
          for (int64_t k = pl.near_ptr[a]; k < pl.near_ptr[a + 1]; ++k)
            if (pl.box_leaf[pl.near_idx[k]]) addbox(pl.near_idx[k]);
        std::sort(rg.begin(), rg.end());
        std::vector<int64_t> merged;
        for (auto& r : rg) {
          if (r.first == r.second) continue;
          if (!merged.empty() && r.first <= merged.back()) {
            if (r.first < merged.back()) overlap = 1;
            merged.back() = std::max(merged.back(), r.second);
          } else {
            merged.push_back(r.first);
            merged.push_back(r.second);
          }
        }
        std::vector<int32_t> cl;
        for (size_t k = 0; k < merged.size(); k += 2)
          for (int64_t i = merged[k]; i < merged[k + 1]; ++i) {
            const int f = pl.face_perm[i];
            for (int c = 0; c < npf; ++c) {
              const int d = face_dof[(int64_t)npf * f + c];
              if (d < 0) continue;
              if (mark[d] != ti) {
                mark[d] = ti;
                cl.push_back(d);
              }
            }
          }
        std::sort(cl.begin(), cl.end());
        ranges[ti] = std::move(merged);
        cols[ti] = std::move(cl);
      }
    }
    if (overlap) throw std::runtime_error("overlapping near ranges");
    BlockGemv& G = pl.near;
    G = BlockGemv();
    pl.near_ncol_dof.clear();
    pl.near_range_ptr.assign(1, 0);
    pl.near_ranges.clear();
    pl.near_range_prefix.clear();
    pl.near_nnz = 0;
    // blocks of at most kMaxRows rows, emitted in descending cost (rows x columns) so that the GPU starts
    // the largest ones first and the end of the launch is made of small blocks (the boxes at the depth
    // limit hold up to ~10x the median number of targets)
    struct Chunk {
      int64_t cost;
      int ti;
      size_t r0;
    };
    std::vector<Chunk> chunks;
    for (int ti = 0; ti < ntb; ++ti) {
      const size_t nt = box_targets[tboxes[ti]].size();
      const int64_t ncol = pad_cols((int)cols[ti].size() + C);
      for (size_t r0 = 0; r0 < nt; r0 += kMaxRows)
        chunks.push_back({(int64_t)(std::min(nt, r0 + (size_t)kMaxRows) - r0) * ncol, ti, r0});
    }
    std::stable_sort(chunks.begin(), chunks.end(), [](const Chunk& a, const Chunk& b) { return a.cost > b.cost; });
    // offsets of every block's rows, columns, matrix and ranges (prefix sums in block order), then the
    // arrays are filled in parallel
    const int nch = (int)chunks.size();
    G.row_ptr.assign(nch + 1, 0);
    G.col_ptr.assign(nch + 1, 0);
    G.mat_ptr.assign(nch, 0);
    pl.near_ncol_dof.assign(nch, 0);
    pl.near_range_ptr.assign(nch + 1, 0);
    G.mat_size = 0;
    for (int k = 0; k < nch; ++k) {
      const Chunk& ch = chunks[k];
      const size_t nt = box_targets[tboxes[ch.ti]].size();
      const int64_t rows = (int64_t)(std::min(nt, ch.r0 + (size_t)kMaxRows) - ch.r0);
      const int ncd = (int)cols[ch.ti].size();
      const int ncol = pad_cols(ncd + C);
      G.row_ptr[k + 1] = G.row_ptr[k] + rows;
      G.col_ptr[k + 1] = G.col_ptr[k] + ncol;
      G.mat_ptr[k] = G.mat_size;
      G.mat_size += rows * 2 * ncol;
      pl.near_ncol_dof[k] = ncd;
      pl.near_range_ptr[k + 1] = pl.near_range_ptr[k] + (int64_t)ranges[ch.ti].size() / 2;
      pl.near_nnz += rows * ncd;
    }
    G.row_out.assign(G.row_ptr[nch], 0);
    G.col_idx.assign(G.col_ptr[nch], 0);
    pl.near_ranges.assign(2 * pl.near_range_ptr[nch], 0);
    pl.near_range_prefix.assign(pl.near_range_ptr[nch], 0);
#pragma omp parallel for schedule(dynamic, 16)
    for (int k = 0; k < nch; ++k) {
      const Chunk& ch = chunks[k];
      const int ti = ch.ti;
      const int b = tboxes[ti];
      const auto& tg = box_targets[b];
      const int ncd = (int)cols[ti].size();
      int64_t o = G.row_ptr[k];
      for (int64_t r = (int64_t)ch.r0; o < G.row_ptr[k + 1]; ++r) G.row_out[o++] = tg[r];
      int32_t* ci = &G.col_idx[G.col_ptr[k]];
      std::copy(cols[ti].begin(), cols[ti].end(), ci);
      for (int c = 0; c < C; ++c) ci[ncd + c] = n_dof + b * C + c;
      for (int64_t c = ncd + C; c < G.col_ptr[k + 1] - G.col_ptr[k]; ++c) ci[c] = n_dof + b * C + C - 1;  // zero-weight padding
      int64_t pre = 0, q = pl.near_range_ptr[k];
      for (size_t j = 0; j < ranges[ti].size(); j += 2, ++q) {
        pl.near_ranges[2 * q] = ranges[ti][j];
        pl.near_ranges[2 * q + 1] = ranges[ti][j + 1];
        pl.near_range_prefix[q] = pre;
        pre += ranges[ti][j + 1] - ranges[ti][j];
      }
    }
    // L2P coefficients per near row: far = (1/s) Re sum R_n^m(x~) L~_n^m, x~ = (x - c)/s
    const int64_t nrows = (int64_t)G.row_out.size();
    pl.l2p_coef.assign((size_t)nrows * C, 0.);
#pragma omp parallel
    {
      Harm H(P);
#pragma omp for schedule(static)
      for (int64_t r = 0; r < nrows; ++r) {
        const int node = G.row_out[r];
        const int b = pl.target_box[node];
        const double s = pl.level_side[pl.box_level[b]];
        const double x = (pl.Xs[3 * node] - pl.box_center[3 * b]) / s;
        const double y = (pl.Xs[3 * node + 1] - pl.box_center[3 * b + 1]) / s;
        const double z = (pl.Xs[3 * node + 2] - pl.box_center[3 * b + 2]) / s;
        H.eval(x, y, z, P);
        double* out = &pl.l2p_coef[(size_t)r * C];
        for (int n = 0; n <= P; ++n)
          for (int m = 0; m <= n; ++m) {
            const double c = H.cnm[Harm::t(n, m)];
            const double Rr = c * H.re[Harm::t(n, m)], Ri = c * H.im[Harm::t(n, m)];
            if (m == 0) {
              out[pk_re(n, 0)] = Rr / s;
            } else {
              out[pk_re(n, m)] = 2.0 * Rr / s;
              out[pk_im(n, m)] = -2.0 * Ri / s;
            }
          }
      }
    }
  }
  if (near_on_host) pl.t_near_struct = now() - t2;

  // --- P2M blocks: one per leaf; M~_n^m = sum W b_k (phi_n, phi) (conj R_n^m(x~), conj dR_n^m.n/s) ---
  const double t3 = now();
  {
    std::vector<int> leaves;
    for (int i = 0; i < nb; ++i)
      if (pl.box_leaf[i]) leaves.push_back(i);
    const int nl = (int)leaves.size();
    std::vector<std::vector<int32_t>> cols(nl);
#pragma omp parallel for schedule(dynamic, 32)
    for (int li = 0; li < nl; ++li) {
      const int b = leaves[li];
      std::vector<int32_t> cl;
      for (int64_t i = pl.box_range[2 * b]; i < pl.box_range[2 * b + 1]; ++i) {
        const int f = pl.face_perm[i];
        for (int c = 0; c < npf; ++c)
          if (face_dof[(int64_t)npf * f + c] >= 0) cl.push_back(face_dof[(int64_t)npf * f + c]);
      }
      std::sort(cl.begin(), cl.end());
      cl.erase(std::unique(cl.begin(), cl.end()), cl.end());
      cols[li] = std::move(cl);
    }
    BlockGemv& G = pl.p2m;
    G = BlockGemv();
    const int CP = (C + 63) / 64 * 64;
    pl.p2m_rows_pad = CP;
    // offsets per leaf (prefix sums), then every leaf's entries filled in parallel
    G.mat_ptr.assign(nl, 0);
    G.row_ptr.assign(nl + 1, 0);
    G.col_ptr.assign(nl + 1, 0);
    pl.p2m_face_ptr.assign(nl + 1, 0);
    for (int li = 0; li < nl; ++li) {
      const int b = leaves[li];
      G.mat_ptr[li] = G.mat_size;
      G.mat_size += (int64_t)CP * 2 * (int64_t)cols[li].size();
      G.row_ptr[li + 1] = G.row_ptr[li] + C;
      G.col_ptr[li + 1] = G.col_ptr[li] + (int64_t)cols[li].size();
      pl.p2m_face_ptr[li + 1] = pl.p2m_face_ptr[li] + (pl.box_range[2 * b + 1] - pl.box_range[2 * b]);
    }
    G.row_out.assign(G.row_ptr[nl], 0);
    G.col_idx.assign(G.col_ptr[nl], 0);
    pl.p2m_faces.assign(pl.p2m_face_ptr[nl], 0);
    pl.p2m_face_lc.assign((size_t)npf * pl.p2m_face_ptr[nl], 0);
    pl.p2m_leaf_cs.assign(4 * (size_t)nl, 0.);
#pragma omp parallel for schedule(dynamic, 32)
    for (int li = 0; li < nl; ++li) {
      const int b = leaves[li];
      for (int r = 0; r < C; ++r) G.row_out[G.row_ptr[li] + r] = b * C + r;
      std::copy(cols[li].begin(), cols[li].end(), &G.col_idx[G.col_ptr[li]]);
      int64_t k = pl.p2m_face_ptr[li];
      for (int64_t i = pl.box_range[2 * b]; i < pl.box_range[2 * b + 1]; ++i, ++k) {
        const int f = pl.face_perm[i];
        pl.p2m_faces[k] = f;
        for (int c = 0; c < npf; ++c) {
          const int d = face_dof[(int64_t)npf * f + c];
          pl.p2m_face_lc[(size_t)npf * k + c] =
              d < 0 ? -1 : (int32_t)(std::lower_bound(cols[li].begin(), cols[li].end(), d) - cols[li].begin());
        }
      }
      for (int a = 0; a < 3; ++a) pl.p2m_leaf_cs[4 * (size_t)li + a] = pl.box_center[3 * b + a];
      pl.p2m_leaf_cs[4 * (size_t)li + 3] = pl.level_side[pl.box_level[b]];
    }
    if (opt.fill_p2m && npf != 3) throw std::runtime_error("host P2M matrices are built for linear faces only");
    if (opt.fill_p2m) pl.p2m_mat.assign(G.mat_size, 0.);
    else pl.p2m_mat.clear();
#pragma omp parallel if (opt.fill_p2m)
    if (opt.fill_p2m) {
      Harm H(P + 1);
#pragma omp for schedule(dynamic, 4)
      for (int li = 0; li < nl; ++li) {
        const int b = leaves[li];
        const double s = pl.level_side[pl.box_level[b]];
        const int ncol = (int)cols[li].size();
        double* M = &pl.p2m_mat[G.mat_ptr[li]];
        for (int64_t fi = pl.box_range[2 * b]; fi < pl.box_range[2 * b + 1]; ++fi) {
          const int f = pl.face_perm[fi];
          int lc[3];
          for (int c = 0; c < 3; ++c)
            lc[c] = (int)(std::lower_bound(cols[li].begin(), cols[li].end(), face_dof[3 * f + c]) - cols[li].begin());
          const double J = pl.fJ[f];
          const double nx = pl.fcr[3 * f] / J, ny = pl.fcr[3 * f + 1] / J, nz = pl.fcr[3 * f + 2] / J;
          for (const auto& rq : rule) {
            const double W = J * rq[3];
            double xq[3];
            for (int a = 0; a < 3; ++a)
              xq[a] = rq[0] * pl.fX[9 * f + a] + rq[1] * pl.fX[9 * f + 3 + a] + rq[2] * pl.fX[9 * f + 6 + a];
            H.eval((xq[0] - pl.box_center[3 * b]) / s, (xq[1] - pl.box_center[3 * b + 1]) / s,
                   (xq[2] - pl.box_center[3 * b + 2]) / s, P);
            for (int n = 0; n <= P; ++n)
              for (int m = 0; m <= n; ++m) {
                const double c = H.cnm[Harm::t(n, m)];
                const double Rr = c * H.re[Harm::t(n, m)], Ri = c * H.im[Harm::t(n, m)];
                // grad R_n^m . n = c(n,m) [ -a NR_{n-1}^{m+1} + b NR_{n-1}^{m-1} + nz NR_{n-1}^m ],
                // a = (nx - i ny)/2, b = (nx + i ny)/2
                double dr = 0, di = 0;
                if (n > 0) {
                  double p1r, p1i, m1r, m1i, z0r, z0i;
                  H.nr(n - 1, m + 1, p1r, p1i);
                  H.nr(n - 1, m - 1, m1r, m1i);
                  H.nr(n - 1, m, z0r, z0i);
                  const double ar = 0.5 * nx, ai = -0.5 * ny, br = 0.5 * nx, bi = 0.5 * ny;
                  dr = -(ar * p1r - ai * p1i) + (br * m1r - bi * m1i) + nz * z0r;
                  di = -(ar * p1i + ai * p1r) + (br * m1i + bi * m1r) + nz * z0i;
                  dr *= c / s;
                  di *= c / s;
                }
                for (int k = 0; k < 3; ++k) {
                  const double wk = W * rq[k];
                  const int col = lc[k];
                  const int rr = pk_re(n, m);
                  M[(size_t)col * CP + rr] += wk * Rr;
                  M[(size_t)(ncol + col) * CP + rr] += wk * dr;
                  if (m > 0) {
                    const int ri = pk_im(n, m);
                    M[(size_t)col * CP + ri] += -wk * Ri;
                    M[(size_t)(ncol + col) * CP + ri] += -wk * di;
                  }
                }
              }
          }
        }
      }
    }
  }
  pl.t_p2m = now() - t3;

  // --- translation pairs and operators ---
  const double t4 = now();
  {
    pl.m2m_by_level.assign(pl.n_levels, PairGroups());
    pl.l2l_by_level.assign(pl.n_levels, PairGroups());
    for (int l = 1; l < pl.n_levels; ++l) {
      std::vector<std::vector<std::pair<int, int>>> g(8);
      for (int c = pl.level_ptr[l]; c < pl.level_ptr[l + 1]; ++c) {
        const int p = pl.box_parent[c];
        int oct = 0;
        for (int a = 0; a < 3; ++a)
          if (pl.box_center[3 * c + a] > pl.box_center[3 * p + a]) oct |= 1 << a;
        g[oct].push_back({p, c});
      }
      for (int o = 0; o < 8; ++o) {
        if (g[o].empty()) continue;
        PairGroups& M = pl.m2m_by_level[l];
        PairGroups& Lg = pl.l2l_by_level[l];
        M.op.push_back(o);
        Lg.op.push_back(8 + o);
        for (auto& pr : g[o]) {
          M.dst.push_back(pr.first);
          M.src.push_back(pr.second);
          Lg.dst.push_back(pr.second);
          Lg.src.push_back(pr.first);
        }
        M.pair_ptr.push_back(M.n_pairs());
        Lg.pair_ptr.push_back(Lg.n_pairs());
      }
    }
    // M2L: L(A) += T(offset) M(B) for B in A's M2L list (buckets_for_L2M is the same set by symmetry).
    // Pairs grouped by integer offset, groups in lexicographic offset order, pairs of a group in target
    // order: offsets are packed into codes ((dx+64) * 128 + dy+64) * 128 + dz+64 (monotone in that order)
    // and the pairs counting-sorted by code (stable).
    constexpr int kOffB = 64;
    std::vector<int64_t> base(nb + 1, 0);
    for (int i = 0; i < nb; ++i) base[i + 1] = base[i] + (pl.m2l_ptr[i + 1] - pl.m2l_ptr[i]);
    const int64_t np = base[nb];
    pl.n_m2l_pairs = np;
    std::vector<int32_t> code(np), srcid(np);
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : bad)
    for (int i = 0; i < nb; ++i) {
      const double s = pl.level_side[pl.box_level[i]];
      const double* tX = &pl.box_center[3 * (size_t)i];
      int64_t k = base[i];
      for (int64_t q = pl.m2l_ptr[i]; q < pl.m2l_ptr[i + 1]; ++q) {
        const int src = pl.m2l_idx[q];
        const double* uX = &pl.box_center[3 * (size_t)src];
        int o[3];
        for (int a = 0; a < 3; ++a) {
          const double v = (uX[a] - tX[a]) / s;
          o[a] = (int)std::lround(v);
          if (std::abs(v - o[a]) > 1e-6 || o[a] <= -kOffB || o[a] >= kOffB) bad = 1;
        }
        code[k] = ((o[0] + kOffB) * 2 * kOffB + (o[1] + kOffB)) * 2 * kOffB + (o[2] + kOffB);
        srcid[k] = src;
        ++k;
      }
    }
    if (bad) throw std::runtime_error("M2L pair not on the level lattice (or offset beyond +-63)");
    const int ncode = 8 * kOffB * kOffB * kOffB;
    std::vector<int64_t> cnt(ncode + 1, 0);
    for (int64_t k = 0; k < np; ++k) cnt[code[k] + 1]++;
    pl.m2l = PairGroups();
    pl.m2l_offsets.clear();
    int opid = 16;
    for (int cd = 0; cd < ncode; ++cd) {
      const int64_t n_in = cnt[cd + 1];
      cnt[cd + 1] += cnt[cd];
      if (n_in == 0) continue;
      pl.m2l.op.push_back(opid++);
      pl.m2l_offsets.push_back(cd / (4 * kOffB * kOffB) - kOffB);
      pl.m2l_offsets.push_back((cd / (2 * kOffB)) % (2 * kOffB) - kOffB);
      pl.m2l_offsets.push_back(cd % (2 * kOffB) - kOffB);
      pl.m2l.pair_ptr.push_back(cnt[cd + 1]);
    }
    pl.m2l.dst.assign(np, 0);
    pl.m2l.src.assign(np, 0);
    for (int i = 0; i < nb; ++i)
      for (int64_t k = base[i]; k < base[i + 1]; ++k) {
        const int64_t at = cnt[code[k]]++;
        pl.m2l.dst[at] = i;
        pl.m2l.src[at] = srcid[k];
      }
    pl.n_ops = opid;
    if (opt.materialize_ops) pl.ops.assign((size_t)pl.n_ops * C * C, 0.);
    else pl.ops.clear();
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < pl.n_ops; ++i) {
      const auto k = op_key(pl, i);
      const auto v = cached_operator(P, k[0], k[1], k[2], k[3]);  // computes it if missing
      if (opt.materialize_ops) std::copy(v->begin(), v->end(), &pl.ops[(size_t)i * C * C]);
    }
  }
  pl.t_ops = now() - t4;
}

void warm_operator_cache(int P, const std::vector<std::array<int, 4>>& keys) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t i = 0; i < (int64_t)keys.size(); ++i) (void)cached_operator(P, keys[i][0], keys[i][1], keys[i][2], keys[i][3]);
}

std::shared_ptr<const std::vector<double>> cached_operator(int P, int kind, int a, int b, int c) {
  static std::mutex mu;
  static std::map<std::array<int, 5>, std::shared_ptr<const std::vector<double>>> cache;
  const std::array<int, 5> key{P, kind, a, b, c};
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  const int C = (P + 1) * (P + 1);
  auto v = std::make_shared<std::vector<double>>(kind == 3 ? (size_t)m2l_rotation_size(P) : (size_t)C * C);
  if (kind == 0) m2m_operator(P, a, v->data());
  else if (kind == 1) l2l_operator(P, a, v->data());
  else if (kind == 2) m2l_operator(P, a, b, c, v->data());
  else if (kind == 3) m2l_rotation_operator(P, a, b, c, v->data());
  else m2l_plane_wave_operator(P, kind == 4 ? 10 : 6, a, b, c, v->data());
  std::lock_guard<std::mutex> lk(mu);
  return cache.emplace(key, std::move(v)).first->second;  // a concurrent insert of the same key wins; same values
}

}  // namespace bemfmm
