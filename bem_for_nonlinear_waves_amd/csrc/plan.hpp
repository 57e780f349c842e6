// Host-side FMM plan: the flat, device-ready description of one BVP built from
// the caller's mesh/DOF arrays. Built by plan.cpp (g++/OpenMP, float128
// operators in operators.cpp), uploaded and executed by bemfmm.hip.
#pragma once
#include <array>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "treemath.hpp"

namespace bemfmm {

struct Options {
  int p = 10;             // expansion order (Buckets<...,N> template argument, BEM_solveBVP.hpp:563)
  double theta = 0.25;    // mac_theta (main_time_domain.cpp:63)
  int leaf_max = 50;      // fmm_bucket_max_points (main_time_domain.cpp:54-55)
  int max_level = 7;      // fmm_max_level
  int q = 6;              // p2m_quadrature_points (main_time_domain.cpp:62)
  int npf = 3;            // nodes per face: 3 linear triangles, 6 true quadratic (vertices, then the midpoints of
                          // edges (0,1), (1,2), (2,0); a midpoint's DOF may be -1: no DOF, weight redistributed)
  bool materialize_ops = true;  // fill Plan::ops (else only make sure every operator is in the cache)
  bool fill_p2m = true;         // fill Plan::p2m_mat on the host (else only its layout and the per-leaf source lists)
};

// Box-blocked dense GEMV: for block b, rows [row_ptr[b], row_ptr[b+1]) and
// gathered columns col_idx[col_ptr[b] .. col_ptr[b+1]).  Row r of block b holds
// two interleaved dense rows (A0 then A1, each ncol long) at mat_ptr[b] + 2*ncol*(r-row_ptr[b]).
//   y0[row_out[r]] = A0[r,:] . x0[col_idx], y1[row_out[r]] = A1[r,:] . x1[col_idx]
struct BlockGemv {
  std::vector<int64_t> row_ptr{0};
  std::vector<int64_t> col_ptr{0};
  std::vector<int64_t> mat_ptr;  // per block
  std::vector<int32_t> row_out;
  std::vector<int32_t> col_idx;
  int64_t mat_size = 0;  // doubles
  int n_blocks() const { return (int)mat_ptr.size(); }
};

// Translation pairs grouped by operator: group g uses operator op[g] and pairs
// [pair_ptr[g], pair_ptr[g+1]) of (dst box, src box).
struct PairGroups {
  std::vector<int32_t> op;
  std::vector<int64_t> pair_ptr{0};
  std::vector<int32_t> dst, src;
  int64_t n_pairs() const { return (int64_t)dst.size(); }
};

struct Plan {
  Options opt;
  int P = 10, C = 121;  // C = (P+1)^2 real-field packed coefficients per layer
  int n_nodes = 0, n_faces = 0, n_dof = 0;
  double Lscale = 1.0;      // coordinate scale (bounding-box diagonal)
  double scale = 1.0;       // getScale() of the scaled geometry
  double near_region = 0;   // scale/25
  double root_side = 0;     // side of the level-0 cube
  // scaled geometry
  std::vector<double> Xs;            // n_nodes*3
  int npf = 3;                       // nodes per face (3 or 6)
  std::vector<double> fX;            // n_faces*3*npf node coordinates
  std::vector<double> fcent;         // n_faces*3 (a+b+c)/3
  std::vector<double> fcr;           // n_faces*3 (2*area*normal)
  std::vector<double> fJ;            // n_faces
  std::vector<int32_t> fv;           // n_faces*npf node ids
  std::vector<int32_t> fdof;         // n_faces*npf DOF ids (-1: quadratic edge midpoint without a DOF)
  std::vector<int32_t> dof_kind, dof_point;
  // inputs of the tree: per face the P2M pole (vertex mean with weights 1/3) and the largest distance from it
  // to the face's quadrature points; the root box (cube of half-side 1.05 scale/2 about the bounds' centre)
  std::vector<double> pole;          // n_faces*3
  std::vector<double> maxoff;        // n_faces
  BoxGeom root{};
  // tree (boxes numbered level-major, preorder within a level)
  int n_levels = 0, n_boxes = 0;
  std::vector<int32_t> level_ptr;    // n_levels+1
  std::vector<int32_t> box_level, box_parent, box_leaf;
  std::vector<double> box_center;    // n_boxes*3
  std::vector<double> level_side;    // ideal side per level (root_side/2^l)
  std::vector<int32_t> face_perm;    // DFS order of faces
  std::vector<int64_t> box_range;    // n_boxes*2 range into face_perm
  // interaction lists by box id, ascending: near[near_ptr[b] .. near_ptr[b+1]) are the boxes of b's level
  // that are not far from b (b excluded), m2l[...] the far boxes among the children of b's parent's near boxes
  std::vector<int64_t> near_ptr, m2l_ptr;  // n_boxes+1
  std::vector<int32_t> near_idx, m2l_idx;
  // near field + L2P (one block per target box)
  BlockGemv near;
  std::vector<int32_t> near_ncol_dof;     // per block: DOF columns (then C L2P columns)
  std::vector<int64_t> near_range_ptr{0}; // per block ranges into face_perm
  std::vector<int64_t> near_ranges;       // pairs (start,end)
  std::vector<int64_t> near_range_prefix; // per range: faces before it within its block
  std::vector<int32_t> target_box;        // per node
  std::vector<double> l2p_coef;           // per near row: C values (row order of near.row_out)
  // P2M (one block per leaf): rows, columns and row_out as a BlockGemv, but the matrix of leaf b is stored
  // column-major per layer, p2m_mat[mat_ptr[b] + (layer * ncol + col) * p2m_rows_pad + row], rows padded
  // to p2m_rows_pad (C rounded up to 64), so that consecutive threads (rows) read consecutive words
  BlockGemv p2m;
  std::vector<double> p2m_mat;            // p2m.mat_size (if fill_p2m)
  int p2m_rows_pad = 0;
  // P2M sources per leaf block (in block order): faces p2m_faces[p2m_face_ptr[l] .. p2m_face_ptr[l+1]), the
  // block-local columns of each face's 3 DOFs (p2m_face_lc, 3 per face entry), and the leaf's centre and
  // side (p2m_leaf_cs, 4 per leaf)
  std::vector<int64_t> p2m_face_ptr;
  std::vector<int32_t> p2m_faces, p2m_face_lc;
  std::vector<double> p2m_leaf_cs;
  // translations
  std::vector<PairGroups> m2m_by_level;   // index = child level
  std::vector<PairGroups> l2l_by_level;   // index = child level
  PairGroups m2l;
  int n_ops = 0;
  std::vector<double> ops;                // n_ops * C * C, layout [op][k_in][r_out] (if materialize_ops)
  std::vector<int32_t> m2l_offsets;       // per M2L op: (dx,dy,dz)
  // stats
  int64_t n_m2l_pairs = 0, near_nnz = 0, n_leaves = 0;
  double t_tree = 0, t_lists = 0, t_near_struct = 0, t_p2m = 0, t_ops = 0;
};

// Builds everything except the near-field quadrature weights (computed on the GPU): plan_geometry, then
// plan_tree_lists, then plan_finish. Throws std::runtime_error on invalid input.
void build_plan(Plan& pl, const Options& opt, int n_nodes, const double* X, int n_faces, const int32_t* tri,
                const int32_t* face_dof, int n_dof, const int32_t* dof_kind, const int32_t* dof_point);
// input checks, coordinate scaling, faces and their P2M sources, the root box
void plan_geometry(Plan& pl, const Options& opt, int n_nodes, const double* X, int n_faces, const int32_t* tri,
                   const int32_t* face_dof, int n_dof, const int32_t* dof_kind, const int32_t* dof_point);
// the octree, its interaction lists and the target box of every node, on the host (sets n_levels .. m2l_idx,
// target_box, n_leaves)
void plan_tree_lists(Plan& pl);
// the box side of every level (from the tree's level count and the root side)
void plan_level_sides(Plan& pl);
// near blocks and L2P coefficients (unless near_on_host is false: the device builds them), P2M blocks and the
// translation groups from the tree arrays; writes no field that the device near-structure stage reads or writes
// (level_side is only set here when plan_level_sides has not run), so the two may run at the same time
void plan_finish(Plan& pl, bool near_on_host = true);

// Real-field packed translation operators (operators.cpp), computed in __float128
// and rounded once to double. Layout: out[k_in * C + r_out].
void m2l_operator(int P, int dx, int dy, int dz, double* out);
// M2L factored as H^T T(|o| z) D (rotation of the offset onto the z axis): out holds, row-major
// [out][in], the degree blocks D_n ((2n+1)^2 each, n = 0..P), the axial blocks ((P+1-m)^2 each,
// m = 0..P, indexed [j-m][n-m]) and the degree blocks of H^T; m2l_rotation_size(P) doubles.
int m2l_rotation_size(int P);
// M2L composed through plane waves (rotation of the dominant axis of the translation onto +z, multipole to
// exponentials, diagonal translation, exponentials to local, rotation back) with the 6- or 10-digit rule,
// in the layout of m2l_operator
void m2l_plane_wave_operator(int P, int digits, int dx, int dy, int dz, double* out);
void m2l_rotation_operator(int P, int dx, int dy, int dz, double* out);
void m2m_operator(int P, int octant, double* out);  // child -> parent; octant bit a set = child on + side of axis a
void l2l_operator(int P, int octant, double* out);  // parent -> child

// Process-wide cache of translation operators (computed once, immutable afterwards):
// kind 0 = M2M of octant a, 1 = L2L of octant a, 2 = M2L of lattice offset (a, b, c), 3 = the rotation
// factors of that M2L (m2l_rotation_operator), 4 / 5 = that M2L through plane waves with the 10- / 6-digit
// rule (m2l_plane_wave_operator). Thread-safe.
std::shared_ptr<const std::vector<double>> cached_operator(int P, int kind, int a, int b, int c);
// computes the missing cache entries of the given (kind, a, b, c) keys in parallel (OpenMP)
void warm_operator_cache(int P, const std::vector<std::array<int, 4>>& keys);
// (kind, a, b, c) of operator `op` of a plan: 0-7 M2M, 8-15 L2L, 16+ M2L offsets
inline std::array<int, 4> op_key(const Plan& pl, int op) {
  if (op < 8) return {0, op, 0, 0};
  if (op < 16) return {1, op - 8, 0, 0};
  const int g = op - 16;
  return {2, pl.m2l_offsets[3 * g], pl.m2l_offsets[3 * g + 1], pl.m2l_offsets[3 * g + 2]};
}

inline int pk_re(int n, int m) { return m == 0 ? n * n : n * n + 2 * m - 1; }
inline int pk_im(int n, int m) { return n * n + 2 * m; }

}  // namespace bemfmm

namespace bemfmm {

// Overlapping DOF blocks of the additive Schwarz preconditioner (the reference's
// SchwarzAdditive_CRS, basic_linear_systems.hpp:2870-3003, blocks built as in
// BEM_solveBVP_GMRES_FMM.hpp:925-1010). Block b holds DOFs idx[ptr[b] .. ptr[b+1]);
// its first ncore[b] entries are its core: every DOF lies in exactly one core.
struct SchwarzBlocks {
  std::vector<int64_t> ptr{0};
  std::vector<int32_t> idx;
  std::vector<int32_t> ncore;
  int max_block = 0;
  int n_blocks() const { return (int)ncore.size(); }
};

// Greedy partition of the DOF graph (DOFs sharing a face or a point are adjacent) into
// breadth-first cores of at most max_core DOFs, seeded in the order of `seed_order`,
// each grown by overlap_k rings of neighbours to at most max_block DOFs.
void build_schwarz_blocks(int n_dof, int n_faces, const int32_t* face_dof, const int32_t* dof_point,
                          const std::vector<int32_t>& seed_order, int max_core, int max_block, int overlap_k,
                          SchwarzBlocks& out, int npf = 3);

}  // namespace bemfmm

namespace bemfmm {

// Sparse rows of the reference's preconditioner matrix (buildSparseMatrixForILU,
// BEM_solveBVP_GMRES_FMM.hpp:215-662): row d over columns col[ptr[d] .. ptr[d+1]) (ascending), diag[d] = position
// of the diagonal entry. After ilu0_factor the strictly lower part holds L (unit diagonal implied) and the
// diagonal and upper part hold U (ILU0_CRS, basic_linear_systems.hpp:2337-2560).
struct SparseRows {
  int n = 0;
  std::vector<int64_t> ptr{0};
  std::vector<int32_t> col;
  std::vector<double> val;
  std::vector<int64_t> diag;
};
// rows of a sweep's dependency levels: rows[ptr[l] .. ptr[l+1]) can be solved once levels < l are done
struct SweepLevels {
  std::vector<int32_t> rows;
  std::vector<int32_t> ptr{0};
  int n_levels() const { return (int)ptr.size() - 1; }
};
// neighborhood 0 = BUCKETS (the reference default), 1 = K-RING with `kring` rings; X are the scaled node
// coordinates, alpha the rigid-mode diagonal per DOF, scale the getScale() of the scaled geometry
void build_ilu_rows(int n_nodes, const double* X, int n_faces, const int32_t* fv, const int32_t* face_dof, int n_dof,
                    const int32_t* dof_kind, const int32_t* dof_point, const double* alpha, double scale,
                    int neighborhood, int kring, SparseRows& A);
// ILU(0) in place (modified: MILU(0), the dropped fill-in times omega moved to the diagonal); throws
// std::runtime_error on a bad pivot
void ilu0_factor(SparseRows& A, double pivot_min = 1e-20, bool modified = false, double omega = 1.0);
// ILUT (threshold dropping, at most max_entries per L and U row): factors of A into F (same row layout)
void ilut_factor(const SparseRows& A, double drop_tol, int max_entries, double pivot_min, SparseRows& F);
void sweep_levels(const SparseRows& A, SweepLevels& forward, SweepLevels& backward);
// the reference's Schwarz blocks on the rows' sparsity graph (cores by breadth-first search to depth core_k)
void build_schwarz_blocks_rows(const SparseRows& A, int core_k, int overlap_k, int max_core, int max_block,
                               SchwarzBlocks& out);

}  // namespace bemfmm
