// The reference's sparse preconditioner matrix and its ILU(0) factorisation (host side).
//
// build_ilu_rows restates buildSparseMatrixForILU for linear elements (bem/BEM_solveBVP_GMRES_FMM.hpp:215-640):
// for every boundary point, the faces of its neighbourhood (BUCKETS: the faces inserted into the point's cell of
// a uniform grid, each face at its centroid, vertices and edge midpoints, plus the point's own faces; K-RING:
// the point's faces grown by k rings of edge neighbours) are integrated with the per-face rules of
// networkFace::setIntegrationInfo (lib/src/networkFace.cpp:28-85): the GW5x5 rule collapsed at the face's
// first point (getPoints(origin), Network.hpp:2062) when the centroid is within scale/10 of the point, else
// GW1x1; on a face that has the point as a vertex only the single-layer weights, GW5x5 collapsed at the point.
// Column of (vertex, face) = that vertex's DOF on the face; its value is G for a Dirichlet column (unknown
// phi_n) and -G_n for the others (unknown phi). Every DOF of the point gets these coefficients; Neumann-type
// rows add -alpha on the diagonal, CORNER rows are the identity, a missing diagonal is set to 1 (:640-662).
//
// build_schwarz_blocks_rows restates the reference's SCHWARZ block partition on the same matrix's graph.
//
// ilu0_factor and sweep_levels restate ILU0_CRS (lib/include/basic_linear_systems.hpp:2337-2560): sorted rows,
// dependency levels of the forward and backward sweeps, IKJ elimination restricted to the pattern (MILU(0): the
// dropped fill-in times omega subtracted from the diagonal), a pivot below 1e-20 is an error; ilut_factor
// restates ILUT_CRS.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "plan.hpp"

namespace bemfmm {

namespace {

// GW5x5 product rule on the unit square (t0, t1, w) (basic_constants.hpp:1326-1350) and GW1x1 (:1306)
const double kGW5R[25][3] = {
    {0.046910077030668, 0.046910077030668, 0.0140335872156063}, {0.046910077030668, 0.230765344947158, 0.0283499999999991},
    {0.046910077030668, 0.5, 0.0336962680968792},               {0.046910077030668, 0.769234655052841, 0.0283499999999991},
    {0.046910077030668, 0.953089922969332, 0.0140335872156063}, {0.230765344947158, 0.046910077030668, 0.0283499999999991},
    {0.230765344947158, 0.230765344947158, 0.0572713510559978}, {0.230765344947158, 0.5, 0.0680716331376877},
    {0.230765344947158, 0.769234655052841, 0.0572713510559978}, {0.230765344947158, 0.953089922969332, 0.0283499999999991},
    {0.5, 0.046910077030668, 0.0336962680968792},               {0.5, 0.230765344947158, 0.0680716331376877},
    {0.5, 0.5, 0.0809086419753086},                             {0.5, 0.769234655052841, 0.0680716331376877},
    {0.5, 0.953089922969332, 0.0336962680968792},               {0.769234655052841, 0.046910077030668, 0.0283499999999991},
    {0.769234655052841, 0.230765344947158, 0.0572713510559978}, {0.769234655052841, 0.5, 0.0680716331376877},
    {0.769234655052841, 0.769234655052841, 0.0572713510559978}, {0.769234655052841, 0.953089922969332, 0.0283499999999991},
    {0.953089922969332, 0.046910077030668, 0.0140335872156063}, {0.953089922969332, 0.230765344947158, 0.0283499999999991},
    {0.953089922969332, 0.5, 0.0336962680968792},               {0.953089922969332, 0.769234655052841, 0.0283499999999991},
    {0.953089922969332, 0.953089922969332, 0.0140335872156063}};
const double kGW1R[1][3] = {{0.5, 0.5, 1.0}};

// single- and double-layer weights of face f (corners in the rotated order `ord`) seen from point x: rule
// collapsed at corner ord[0]; singular (x is corner ord[0]): single layer only
void face_rule(const double* X, const int32_t* v, const int ord[3], const double* x, bool singular, bool near,
               double G[3], double H[3]) {
  const double* P0 = X + 3 * v[ord[0]];
  const double* P1 = X + 3 * v[ord[1]];
  const double* P2 = X + 3 * v[ord[2]];
  const double e1[3] = {P1[0] - P0[0], P1[1] - P0[1], P1[2] - P0[2]};
  const double e2[3] = {P2[0] - P0[0], P2[1] - P0[1], P2[2] - P0[2]};
  const double cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
  const double J = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
  G[0] = G[1] = G[2] = H[0] = H[1] = H[2] = 0.;
  const bool gw5 = singular || near;
  const int nq = gw5 ? 25 : 1;
  for (int g = 0; g < nq; ++g) {
    const double t0 = gw5 ? kGW5R[g][0] : kGW1R[g][0], t1 = gw5 ? kGW5R[g][1] : kGW1R[g][1];
    const double ww = gw5 ? kGW5R[g][2] : kGW1R[g][2];
    const double N[3] = {t0, t1 * (1. - t0), (t0 - 1.) * (t1 - 1.)};
    const double w = ww * (1. - t0);
    double R[3];
    for (int a = 0; a < 3; ++a) R[a] = N[0] * P0[a] + N[1] * P1[a] + N[2] * P2[a] - x[a];
    const double nr = std::sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
    const double ig = J * (w / nr);
    const double ign = (R[0] * cr[0] + R[1] * cr[1] + R[2] * cr[2]) * (w / nr) / (nr * nr);
    for (int k = 0; k < 3; ++k) {
      G[k] += ig * N[k];
      if (!singular) H[k] -= ign * N[k];
    }
  }
}

}  // namespace

void build_ilu_rows(int n_nodes, const double* X, int n_faces, const int32_t* fv, const int32_t* face_dof, int n_dof,
                    const int32_t* dof_kind, const int32_t* dof_point, const double* alpha, double scale,
                    int neighborhood, int kring, SparseRows& A) {
  if (neighborhood < 0 || kring < 0 || kring > 20) throw std::runtime_error("ILU matrix: invalid neighbourhood");
  // faces of each point
  std::vector<int64_t> pf_ptr(n_nodes + 1, 0);
  for (int f = 0; f < n_faces; ++f)
    for (int a = 0; a < 3; ++a) pf_ptr[fv[3 * f + a] + 1]++;
  for (int i = 0; i < n_nodes; ++i) pf_ptr[i + 1] += pf_ptr[i];
  std::vector<int32_t> pf(pf_ptr[n_nodes]);
  {
    std::vector<int64_t> fill(pf_ptr.begin(), pf_ptr.end() - 1);
    for (int f = 0; f < n_faces; ++f)
      for (int a = 0; a < 3; ++a) pf[fill[fv[3 * f + a]]++] = f;
  }
  // DOFs of each point
  std::vector<int64_t> pd_ptr(n_nodes + 1, 0);
  for (int d = 0; d < n_dof; ++d) pd_ptr[dof_point[d] + 1]++;
  for (int i = 0; i < n_nodes; ++i) pd_ptr[i + 1] += pd_ptr[i];
  std::vector<int32_t> pd(n_dof);
  {
    std::vector<int64_t> fill(pd_ptr.begin(), pd_ptr.end() - 1);
    for (int d = 0; d < n_dof; ++d) pd[fill[dof_point[d]]++] = d;
  }
  // BUCKETS: a uniform grid over the cube of half-side 1.05 max-extent around the bounding-box centre, with
  // clamp(round(cbrt(faces)), 8, 64) cells per axis of its side; faces inserted at 7 points each
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int i = 0; i < n_nodes; ++i)
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], X[3 * i + a]);
      hi[a] = std::max(hi[a], X[3 * i + a]);
    }
  double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]});
  double gmin[3];
  for (int a = 0; a < 3; ++a) gmin[a] = 0.5 * (lo[a] + hi[a]) - 1.05 * ext;
  const double side = 2.0 * 1.05 * ext;
  const int cells = std::clamp((int)std::llround(std::cbrt((double)std::max(n_faces, 1))), 8, 64);
  const double dL = side > 0. ? side / cells : 1.0, inv_dL = 1.0 / dL;
  int gs[3];
  for (int a = 0; a < 3; ++a) {
    const double hiA = 0.5 * (lo[a] + hi[a]) + 1.05 * ext;
    gs[a] = std::max(1, (int)std::ceil((hiA - gmin[a]) / dL));
  }
  auto cell_of = [&](const double* x) {
    int64_t id = 0;
    for (int a = 0; a < 3; ++a) {
      int i = (int)std::floor((x[a] - gmin[a]) * inv_dL);
      i = std::clamp(i, 0, gs[a] - 1);
      id = id * gs[a] + i;
    }
    return id;
  };
  std::vector<std::pair<int64_t, int32_t>> cell_face;
  if (neighborhood == 0) {
    cell_face.reserve((size_t)n_faces * 7);
    for (int f = 0; f < n_faces; ++f) {
      const double* P[3] = {X + 3 * fv[3 * f], X + 3 * fv[3 * f + 1], X + 3 * fv[3 * f + 2]};
      double pts[7][3];
      for (int a = 0; a < 3; ++a) {
        pts[0][a] = (P[0][a] + P[1][a] + P[2][a]) / 3.;
        pts[1][a] = P[0][a];
        pts[2][a] = P[1][a];
        pts[3][a] = P[2][a];
        pts[4][a] = (P[0][a] + P[1][a]) / 2.;
        pts[5][a] = (P[1][a] + P[2][a]) / 2.;
        pts[6][a] = (P[2][a] + P[0][a]) / 2.;
      }
      for (auto& q : pts) cell_face.push_back({cell_of(q), f});
    }
    std::sort(cell_face.begin(), cell_face.end());
    cell_face.erase(std::unique(cell_face.begin(), cell_face.end()), cell_face.end());
  }
  // K-RING: edge neighbours of each face
  std::vector<int32_t> fnb;
  if (neighborhood == 1) {
    std::vector<std::pair<int64_t, int32_t>> edges;
    edges.reserve((size_t)n_faces * 3);
    for (int f = 0; f < n_faces; ++f)
      for (int a = 0; a < 3; ++a) {
        int64_t u = fv[3 * f + a], v = fv[3 * f + (a + 1) % 3];
        if (u > v) std::swap(u, v);
        edges.push_back({u * n_nodes + v, f});
      }
    std::sort(edges.begin(), edges.end());
    fnb.assign((size_t)n_faces * 3, -1);
    std::vector<int> nfill(n_faces, 0);
    for (size_t i = 0; i + 1 < edges.size(); ++i)
      if (edges[i].first == edges[i + 1].first) {
        const int32_t f = edges[i].second, g = edges[i + 1].second;
        if (nfill[f] < 3) fnb[3 * f + nfill[f]++] = g;
        if (nfill[g] < 3) fnb[3 * g + nfill[g]++] = f;
      }
  }
  const double near_switch = scale / 10.;
  std::vector<std::vector<std::pair<int32_t, double>>> rows(n_dof);
#pragma omp parallel
  {
    std::vector<int32_t> faces, front, next;
    std::vector<std::pair<int32_t, double>> coef;
#pragma omp for schedule(dynamic, 256)
    for (int k = 0; k < n_nodes; ++k) {
      if (pd_ptr[k] == pd_ptr[k + 1]) continue;
      faces.assign(pf.begin() + pf_ptr[k], pf.begin() + pf_ptr[k + 1]);
      if (neighborhood == 0) {
        const int64_t c = cell_of(X + 3 * k);
        auto it = std::lower_bound(cell_face.begin(), cell_face.end(), std::make_pair(c, (int32_t)-1));
        for (; it != cell_face.end() && it->first == c; ++it) faces.push_back(it->second);
      } else {
        std::sort(faces.begin(), faces.end());
        front = faces;
        for (int r = 0; r < kring; ++r) {
          next.clear();
          for (int32_t f : front)
            for (int a = 0; a < 3; ++a) {
              const int32_t g = fnb[3 * f + a];
              if (g >= 0 && !std::binary_search(faces.begin(), faces.end(), g)) next.push_back(g);
            }
          std::sort(next.begin(), next.end());
          next.erase(std::unique(next.begin(), next.end()), next.end());
          std::vector<int32_t> merged;
          std::merge(faces.begin(), faces.end(), next.begin(), next.end(), std::back_inserter(merged));
          faces.swap(merged);
          front.swap(next);
        }
      }
      std::sort(faces.begin(), faces.end());
      faces.erase(std::unique(faces.begin(), faces.end()), faces.end());
      coef.clear();
      const double* x = X + 3 * k;
      for (int32_t f : faces) {
        const int32_t* v = fv + 3 * f;
        int ord[3];
        if (v[0] == k) ord[0] = 0, ord[1] = 1, ord[2] = 2;
        else if (v[1] == k) ord[0] = 1, ord[1] = 2, ord[2] = 0;
        else ord[0] = 2, ord[1] = 0, ord[2] = 1;
        const bool singular = v[ord[0]] == k;
        double cx[3];
        for (int a = 0; a < 3; ++a) cx[a] = (X[3 * v[0] + a] + X[3 * v[1] + a] + X[3 * v[2] + a]) / 3. - x[a];
        const bool near = std::sqrt(cx[0] * cx[0] + cx[1] * cx[1] + cx[2] * cx[2]) < near_switch;
        double G[3], H[3];
        face_rule(X, v, ord, x, singular, near, G, H);
        for (int a = 0; a < 3; ++a) {
          const int32_t col = face_dof[3 * f + ord[a]];
          const double val = dof_kind[col] == 0 ? G[a] : -H[a];
          if (val != 0.) coef.push_back({col, val});
        }
      }
      // merge duplicate columns (summed in the order above)
      std::stable_sort(coef.begin(), coef.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
      std::vector<std::pair<int32_t, double>> merged;
      for (const auto& e : coef) {
        if (!merged.empty() && merged.back().first == e.first) merged.back().second += e.second;
        else merged.push_back(e);
      }
      for (int64_t q = pd_ptr[k]; q < pd_ptr[k + 1]; ++q) {
        const int32_t d = pd[q];
        auto& row = rows[d];
        if (dof_kind[d] == 2) {  // CORNER Neumann-face DOF: phi = phi_D
          row.assign(1, {d, 1.0});
          continue;
        }
        row = merged;
        auto it = std::lower_bound(row.begin(), row.end(), std::make_pair(d, -1e300));
        const bool has = it != row.end() && it->first == d;
        if (dof_kind[d] != 0) {  // unknown phi: the jump term
          if (has) it->second -= alpha[d];
          else row.insert(it, {d, -alpha[d]});
        } else if (!has) {
          row.insert(it, {d, 1.0});
        }
      }
    }
  }
  A.n = n_dof;
  A.ptr.assign(n_dof + 1, 0);
  for (int d = 0; d < n_dof; ++d) A.ptr[d + 1] = A.ptr[d] + (int64_t)rows[d].size();
  A.col.resize(A.ptr[n_dof]);
  A.val.resize(A.ptr[n_dof]);
  A.diag.assign(n_dof, -1);
  for (int d = 0; d < n_dof; ++d) {
    if (rows[d].empty()) throw std::runtime_error("ILU matrix: DOF " + std::to_string(d) + " has no row");
    int64_t q = A.ptr[d];
    for (const auto& e : rows[d]) {
      if (e.first == d) A.diag[d] = q;
      A.col[q] = e.first;
      A.val[q] = e.second;
      ++q;
    }
    if (A.diag[d] < 0) throw std::runtime_error("ILU matrix: missing diagonal at row " + std::to_string(d));
  }
}

void ilu0_factor(SparseRows& A, double pivot_min, bool modified, double omega) {
  const int n = A.n;
  if (modified && !(omega >= 0. && std::isfinite(omega))) throw std::runtime_error("MILU(0): omega must be >= 0");
  for (int i = 0; i < n; ++i) {
    const int64_t di = A.diag[i], end = A.ptr[i + 1];
    for (int64_t ik = A.ptr[i]; ik < di; ++ik) {
      const int k = A.col[ik];
      const double piv = A.val[A.diag[k]];
      if (!std::isfinite(piv) || std::abs(piv) < pivot_min)
        throw std::runtime_error("ILU(0): bad pivot at row " + std::to_string(k) + " while factoring row " +
                                 std::to_string(i));
      A.val[ik] /= piv;
      const double lik = A.val[ik];
      int64_t pi = ik + 1, pk = A.diag[k] + 1;
      const int64_t kend = A.ptr[k + 1];
      while (pi < end && pk < kend) {
        const int ci = A.col[pi], ck = A.col[pk];
        if (ci == ck) {
          A.val[pi] -= lik * A.val[pk];
          ++pi;
          ++pk;
        } else if (ci < ck) {
          ++pi;
        } else {
          // fill-in outside the pattern: dropped (MILU(0) moves it to the diagonal)
          if (modified) A.val[di] -= omega * (lik * A.val[pk]);
          ++pk;
        }
      }
      if (modified)
        for (; pk < kend; ++pk) A.val[di] -= omega * (lik * A.val[pk]);
    }
  }
  for (int i = 0; i < n; ++i) {
    const double d = A.val[A.diag[i]];
    if (!std::isfinite(d) || std::abs(d) < pivot_min)
      throw std::runtime_error("ILU(0): zero or non-finite pivot at row " + std::to_string(i));
  }
}

// ILUT_CRS (basic_linear_systems.hpp:2610-2760): row by row, the working row eliminated by the finished U rows
// with fill-in allowed; entries below drop_tol * |row of A|_1 dropped, then the max_entries largest of the L and
// of the U part kept; a diagonal below pivot_min is set to +-pivot_min. F gets L (strictly lower), the diagonal
// and U in one row each.
void ilut_factor(const SparseRows& A, double drop_tol, int max_entries, double pivot_min, SparseRows& F) {
  const int n = A.n;
  if (max_entries < 0 || !(drop_tol >= 0.) || !(pivot_min > 0.)) throw std::runtime_error("ILUT: invalid parameters");
  std::vector<std::vector<std::pair<int32_t, double>>> U(n);  // diagonal first, then the upper part ascending
  std::vector<std::vector<std::pair<int32_t, double>>> L(n);
  std::vector<double> udiag(n, 0.);
  auto keep_largest = [&](std::vector<std::pair<int32_t, double>>& v) {
    if ((int)v.size() <= max_entries) return;
    std::nth_element(v.begin(), v.begin() + max_entries, v.end(),
                     [](const auto& a, const auto& b) { return std::abs(a.second) > std::abs(b.second); });
    v.resize(max_entries);
  };
  std::map<int32_t, double> w;
  for (int i = 0; i < n; ++i) {
    w.clear();
    double norm1 = 0.;
    for (int64_t q = A.ptr[i]; q < A.ptr[i + 1]; ++q) {
      w[A.col[q]] += A.val[q];
      norm1 += std::abs(A.val[q]);
    }
    const double thr = drop_tol * norm1;
    for (auto it = w.begin(); it != w.end() && it->first < i; ++it) {
      const int k = it->first;
      if (std::abs(udiag[k]) < pivot_min) throw std::runtime_error("ILUT: zero pivot at row " + std::to_string(k));
      const double lik = it->second / udiag[k];
      it->second = lik;
      for (const auto& [j, ukj] : U[k])
        if (j > k) w[j] -= lik * ukj;
    }
    double diag = 0.;
    if (auto it = w.find(i); it != w.end()) diag = it->second;
    if (std::abs(diag) < pivot_min) diag = diag >= 0. ? pivot_min : -pivot_min;
    std::vector<std::pair<int32_t, double>> lc, uc;
    for (const auto& [c, v] : w) {
      if (c == i || v == 0.) continue;
      if (drop_tol > 0. && std::abs(v) < thr) continue;
      (c < i ? lc : uc).push_back({c, v});
    }
    keep_largest(lc);
    keep_largest(uc);
    std::sort(lc.begin(), lc.end());
    std::sort(uc.begin(), uc.end());
    uc.insert(uc.begin(), {i, diag});
    L[i] = std::move(lc);
    U[i] = std::move(uc);
    udiag[i] = diag;
  }
  F = SparseRows{};
  F.n = n;
  F.ptr.assign(n + 1, 0);
  F.diag.assign(n, -1);
  for (int i = 0; i < n; ++i) {
    for (const auto& e : L[i]) {
      F.col.push_back(e.first);
      F.val.push_back(e.second);
    }
    F.diag[i] = (int64_t)F.col.size();
    for (const auto& e : U[i]) {
      F.col.push_back(e.first);
      F.val.push_back(e.second);
    }
    F.ptr[i + 1] = (int64_t)F.col.size();
  }
}

void sweep_levels(const SparseRows& A, SweepLevels& Lv, SweepLevels& Uv) {
  const int n = A.n;
  std::vector<int32_t> lev(n, 0);
  int nl = 0;
  for (int i = 0; i < n; ++i) {
    int m = -1;
    for (int64_t q = A.ptr[i]; q < A.diag[i]; ++q) m = std::max(m, lev[A.col[q]]);
    lev[i] = m + 1;
    nl = std::max(nl, lev[i] + 1);
  }
  auto bucket = [&](SweepLevels& S, int levels) {
    S.ptr.assign(levels + 1, 0);
    for (int i = 0; i < n; ++i) S.ptr[lev[i] + 1]++;
    for (int l = 0; l < levels; ++l) S.ptr[l + 1] += S.ptr[l];
    S.rows.resize(n);
    std::vector<int32_t> fill(S.ptr.begin(), S.ptr.end() - 1);
    for (int i = 0; i < n; ++i) S.rows[fill[lev[i]]++] = i;  // ascending row order within a level
  };
  bucket(Lv, nl);
  nl = 0;
  for (int i = n - 1; i >= 0; --i) {
    int m = -1;
    for (int64_t q = A.diag[i] + 1; q < A.ptr[i + 1]; ++q) m = std::max(m, lev[A.col[q]]);
    lev[i] = m + 1;
    nl = std::max(nl, lev[i] + 1);
  }
  bucket(Uv, nl);
}

// The reference's Schwarz blocks (BEM_solveBVP_GMRES_FMM.hpp:901-1010) on the sparsity graph of the rows
// (i -> its columns j != i, ascending): seeds in row order; a core is a breadth-first search from the seed over
// unassigned DOFs to depth core_k, at most max_core DOFs; the block adds overlap_k rings of neighbours of the
// core (assigned or not) up to max_block DOFs.
void build_schwarz_blocks_rows(const SparseRows& A, int core_k, int overlap_k, int max_core, int max_block,
                               SchwarzBlocks& out) {
  const int n = A.n;
  if (max_core < 1 || max_block < max_core) throw std::runtime_error("schwarz: need 1 <= max_core <= max_block");
  core_k = std::clamp(core_k, 0, 8);
  overlap_k = std::clamp(overlap_k, 0, 8);
  std::vector<char> assigned(n, 0);
  std::vector<int> mark(n, -1);
  int stamp = 0;
  std::vector<std::pair<int, int>> q;  // (DOF, depth), consumed from `head`
  out = SchwarzBlocks{};
  std::vector<int32_t> core, block;
  for (int seed = 0; seed < n; ++seed) {
    if (assigned[seed]) continue;
    core.clear();
    ++stamp;
    q.clear();
    size_t head = 0;
    q.push_back({seed, 0});
    mark[seed] = stamp;
    while (head < q.size() && (int)core.size() < max_core) {
      const auto [u, dep] = q[head++];
      if (!assigned[u]) {
        assigned[u] = 1;
        core.push_back(u);
        if ((int)core.size() >= max_core) break;
      }
      if (dep >= core_k) continue;
      for (int64_t e = A.ptr[u]; e < A.ptr[u + 1]; ++e) {
        const int v = A.col[e];
        if (v == u || mark[v] == stamp) continue;
        mark[v] = stamp;
        if (assigned[v]) continue;
        q.push_back({v, dep + 1});
      }
    }
    if (core.empty()) {
      assigned[seed] = 1;
      core.push_back(seed);
    }
    block = core;
    ++stamp;
    q.clear();
    head = 0;
    for (int u : core) {
      mark[u] = stamp;
      q.push_back({u, 0});
    }
    while (head < q.size() && (int)block.size() < max_block) {
      const auto [u, dep] = q[head++];
      if (dep >= overlap_k) continue;
      for (int64_t e = A.ptr[u]; e < A.ptr[u + 1]; ++e) {
        const int v = A.col[e];
        if (v == u || mark[v] == stamp) continue;
        mark[v] = stamp;
        block.push_back(v);
        if ((int)block.size() >= max_block) break;
        q.push_back({v, dep + 1});
      }
    }
    out.idx.insert(out.idx.end(), block.begin(), block.end());
    out.ptr.push_back((int64_t)out.idx.size());
    out.ncore.push_back((int32_t)core.size());
    out.max_block = std::max(out.max_block, (int)block.size());
  }
}

}  // namespace bemfmm
