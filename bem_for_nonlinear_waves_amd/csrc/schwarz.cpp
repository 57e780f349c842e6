// Overlapping DOF blocks of the additive Schwarz preconditioner.
//
// The reference builds its blocks on the sparsity graph of a near-field sparse
// matrix (BEM_solveBVP_GMRES_FMM.hpp:909-1010): a breadth-first core of at most
// max_core unassigned DOFs around each unassigned seed, then an overlap of
// overlap_k rings of neighbours (assigned or not) up to max_block DOFs. Here the
// graph is the mesh DOF graph (DOFs that share a face or a point), whose rings are
// much thinner than the reference's k-ring matrix pattern, so a core grows to
// max_core DOFs breadth-first without a depth limit. Seeds are taken in the order
// the caller passes (the tree order of the DOFs' points), so cores are compact
// patches of the surface.
#include <algorithm>
#include <deque>
#include <stdexcept>
#include <vector>

#include "plan.hpp"

namespace bemfmm {

void build_schwarz_blocks(int n_dof, int n_faces, const int32_t* face_dof, const int32_t* dof_point,
                          const std::vector<int32_t>& seed_order, int max_core, int max_block, int overlap_k,
                          SchwarzBlocks& out, int npf) {
  if (max_core < 1 || max_block < max_core) throw std::runtime_error("schwarz: need 1 <= max_core <= max_block");
  if ((int)seed_order.size() != n_dof) throw std::runtime_error("schwarz: seed order must list every DOF once");
  // adjacency (CSR) of the DOF graph
  std::vector<std::vector<int32_t>> adj(n_dof);
  for (int f = 0; f < n_faces; ++f)
    for (int a = 0; a < npf; ++a)
      for (int b = 0; b < npf; ++b) {
        const int32_t da = face_dof[(int64_t)npf * f + a], db = face_dof[(int64_t)npf * f + b];
        if (a != b && da >= 0 && db >= 0) adj[da].push_back(db);
      }
  {
    int n_points = 0;
    for (int d = 0; d < n_dof; ++d) n_points = std::max(n_points, dof_point[d] + 1);
    std::vector<std::vector<int32_t>> by_point(n_points);
    for (int d = 0; d < n_dof; ++d) by_point[dof_point[d]].push_back(d);
    for (const auto& ds : by_point)
      for (int32_t a : ds)
        for (int32_t b : ds)
          if (a != b) adj[a].push_back(b);
  }
  for (auto& a : adj) {
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }
  out = SchwarzBlocks{};
  out.max_block = 0;
  std::vector<char> assigned(n_dof, 0);
  std::vector<int> mark(n_dof, -1);
  int stamp = 0;
  std::deque<std::pair<int32_t, int>> q;
  std::vector<int32_t> core, block;
  for (int32_t seed : seed_order) {
    if (seed < 0 || seed >= n_dof) throw std::runtime_error("schwarz: seed out of range");
    if (assigned[seed]) continue;
    // core: breadth-first over unassigned DOFs
    core.clear();
    ++stamp;
    q.clear();
    q.emplace_back(seed, 0);
    mark[seed] = stamp;
    while (!q.empty() && (int)core.size() < max_core) {
      const int32_t u = q.front().first;
      q.pop_front();
      assigned[u] = 1;
      core.push_back(u);
      for (int32_t v : adj[u]) {
        if (mark[v] == stamp || assigned[v]) continue;
        mark[v] = stamp;
        q.emplace_back(v, 0);
      }
    }
    // overlap: overlap_k rings around the core
    block = core;
    ++stamp;
    q.clear();
    for (int32_t u : core) {
      mark[u] = stamp;
      q.emplace_back(u, 0);
    }
    while (!q.empty() && (int)block.size() < max_block) {
      const auto [u, d] = q.front();
      q.pop_front();
      if (d >= overlap_k) continue;
      for (int32_t v : adj[u]) {
        if (mark[v] == stamp) continue;
        mark[v] = stamp;
        block.push_back(v);
        if ((int)block.size() >= max_block) break;
        q.emplace_back(v, d + 1);
      }
    }
    out.idx.insert(out.idx.end(), block.begin(), block.end());
    out.ptr.push_back((int64_t)out.idx.size());
    out.ncore.push_back((int32_t)core.size());
    out.max_block = std::max(out.max_block, (int)block.size());
  }
}

}  // namespace bemfmm
