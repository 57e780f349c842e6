// Launch interface of the device kernels (kernels.hip) for the host side of the library (bemfmm.cpp).
//
// Every function enqueues one kernel (or, for the tiled and templated ones, the instance that matches
// its arguments) on the given HIP stream and returns without synchronising; launch errors surface
// through hipGetLastError as for any HIP launch. Pointers are device memory. The host side includes
// only this header and the HIP runtime API, so it holds no device code.
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>

#include "treemath.hpp"

#ifdef __HIPCC__
#define BEMFMM_HD __host__ __device__
#else
#define BEMFMM_HD
#endif

namespace bemfmm {
namespace dev {

// ---- layout constants shared by the kernels and the host ----
constexpr int kTP = 8;       // pairs per tile of the M2M/L2L translations: their small upper levels are
                             // latency-bound, so more, smaller workgroups pay (209k tank, per matvec:
                             // M2M and L2L each 0.111 ms at 32 pairs, 0.091 at 16, 0.078 at 8)
constexpr int kTPM2L = 16;   // default pairs per M2L tile: 3 workgroups per CU (FP64) instead of 2 at 32 pairs;
                             // measured on the 209k tank: FP64 3.61 -> 3.41 ms, FP32 2.52 -> 2.17 ms; 64 pairs: 4.6 ms
constexpr int kXCD = 8;      // XCDs of the MI355X: workgroup id % 8 selects the XCD
constexpr int kKG = 4;       // k-steps (of 4) per pipeline stage of the matrix-core translation
constexpr int kRotWavePairs = 32;
constexpr int kRotTP = 4 * kRotWavePairs;  // pairs per workgroup tile of the rotation M2L
constexpr int kSchwarzMaxBlock = 128;      // DOFs per Schwarz block (LDS of the block build)
constexpr int kDotBlocks = 256;            // partial sums per dot product

// LDS elements of a translation tile of TP pairs: the B stage (KP x (2TP+1)) or the output transpose
// (2TP x (RP+1))
inline size_t translate_lds(int TP, int KP, int RP) { return (size_t)std::max(KP * (2 * TP + 1), 2 * TP * (RP + 1)); }

// Values per staged M2L column. FP32: C rounded up to 32, so every column starts on a 128-byte line and the
// staging is written and read in whole lines (at stride C = 121 the rotation M2L's 16-row stores straddled
// lines, and its writes and the reduction's reads moved 14% and 7% more than the staged bytes); C <= 128
// (checked in prepare) keeps it <= 128. FP64: C (each wave writes whole 968-byte columns at once: no
// partial lines to remove, and the padding only added bytes).
template <typename T>
BEMFMM_HD constexpr int stage_stride(int C) { return sizeof(T) == 4 ? (C + 31) & ~31 : C; }

// Device layout of the rotation-M2L factors of one offset, in float4s, every row zero-padded to whole
// float4s (the z-rotation pairs are read as float4s; a factor row is read one float per lane, lane l taking
// entry l & 15): (cos m phi, sin m phi) for m = 1..P; per degree n the y-rotation blocks Y_D (cos-type
// (n+1) rows, sin-type n rows); per order m the axial block (P+1-m rows); per degree the blocks of Y_H^T.
// Row lengths equal the block sizes.
BEMFMM_HD constexpr int rot_q(int len) { return (len + 3) / 4; }
BEMFMM_HD constexpr int rot_size4(int P) {
  int s = rot_q(2 * P);
  for (int n = 0; n <= P; ++n) s += 2 * ((n + 1) * rot_q(n + 1) + n * rot_q(n));
  for (int m = 0; m <= P; ++m) s += (P + 1 - m) * rot_q(P + 1 - m);
  return s;
}
BEMFMM_HD constexpr int rot_size(int P) { return 4 * rot_size4(P); }  // floats per offset
// compact factors (m2l_rotation_operator, operators.cpp) -> the padded device layout
inline void rot_pad(int P, const double* in, float* out) {
  std::fill(out, out + rot_size(P), 0.f);
  int io = 0, oo = 0;
  for (int k = 0; k < 2 * P; ++k) out[oo + k] = (float)in[io + k];
  io += 2 * P;
  oo += 4 * rot_q(2 * P);
  auto block = [&](int rows) {
    const int q = 4 * rot_q(rows);
    for (int a = 0; a < rows; ++a)
      for (int c2 = 0; c2 < rows; ++c2) out[oo + a * q + c2] = (float)in[io + a * rows + c2];
    io += rows * rows;
    oo += rows * q;
  };
  for (int n = 0; n <= P; ++n) {
    block(n + 1);
    block(n);
  }
  for (int m = 0; m <= P; ++m) block(P + 1 - m);
  for (int n = 0; n <= P; ++n) {
    block(n + 1);
    block(n);
  }
}

// ---- argument bundles ----
// Box-blocked dense GEMV (near field + L2P, and the P2M leaf blocks): block b has rows
// [row_ptr[b], row_ptr[b+1]) written to row_out, columns col_idx[col_ptr[b] ..), its matrix at mat_ptr[b].
struct Gemv {
  const int64_t* row_ptr = nullptr;
  const int64_t* col_ptr = nullptr;
  const int64_t* mat_ptr = nullptr;
  const int32_t* row_out = nullptr;
  const int32_t* col_idx = nullptr;
  const int32_t* ncol_dof = nullptr;  // near blocks: DOF columns before the L2P columns
  int nblk = 0;
  // launch shape: near GEMV 0 = 4 rows per wave, 1 = 2, 2 = 8; P2M 0 = 8 loads in flight per thread, 1 = 4, 2 = 16
  int variant = 0;
};
// Operator-grouped translation tiles: tile t applies operator slot op[t] to the pairs [p0[t], p1[t]).
struct Tiles {
  const int32_t* op = nullptr;
  const int64_t* p0 = nullptr;
  const int64_t* p1 = nullptr;
  const int32_t* dst = nullptr;
  const int32_t* src = nullptr;
  int ntiles = 0;
  int tp = kTP;  // pairs per tile
};
// Near-field quadrature inputs (k_near_quad)
constexpr int kNearQuadLdsCols = 4096;  // DOF columns of a near row accumulated in LDS (2 x 32 KB)
struct NearQuad {
  int C = 0, q = 0;
  int npf = 3;                        // nodes per face: 3 linear, 6 true quadratic (fX 18, fv / fdof 6 per face)
  const double* dun6_rule = nullptr;  // quadratic faces: the 6-point Dunavant rule of far non-adjacent faces
  int max_ncd = 0;  // largest DOF column count of a near block
  // optional: the local column of each face corner, block b's faces from 3 * block_fstart[b] (else searched)
  const int32_t* face_cols = nullptr;
  const int64_t* block_fstart = nullptr;
  double near_region = 0;
  const int32_t* row_block = nullptr;
  const int32_t* ncol_dof = nullptr;
  const int64_t* range_ptr = nullptr;
  const int64_t* ranges = nullptr;
  const int64_t* range_prefix = nullptr;
  const int32_t* face_perm = nullptr;
  const double* Xs = nullptr;
  const double* fX = nullptr;
  const double* fcent = nullptr;
  const double* fcr = nullptr;
  const double* fJ = nullptr;
  const int32_t* fv = nullptr;
  const int32_t* fdof = nullptr;
  const double* l2p = nullptr;
  const double* dun_rule = nullptr;
};
// Resident near-field rows read by the preconditioner builds (operator entries by DOF)
struct NearRows {
  const int32_t* kind = nullptr;
  const double* kappa = nullptr;
  const int32_t* point = nullptr;
  const int64_t* node_row = nullptr;
  const int32_t* row_block = nullptr;
  const int64_t* row_ptr = nullptr;
  const int64_t* col_ptr = nullptr;
  const int64_t* mat_ptr = nullptr;
  const int32_t* col_idx = nullptr;
  const int32_t* ncol_dof = nullptr;
  const double* mat = nullptr;
  const double* alpha = nullptr;
};
// P2M matrix build inputs (k_p2m_build)
struct P2MBuild {
  int q = 0, CP = 0, nleaves = 0;
  int npf = 3;  // nodes (and DOF columns) per face: 3 linear, 6 true quadratic
  const int64_t* face_ptr = nullptr;
  const int32_t* faces = nullptr;
  const int32_t* face_lc = nullptr;
  const double* leaf_cs = nullptr;
  const int64_t* mat_ptr = nullptr;
  const int64_t* col_ptr = nullptr;
  const double* fX = nullptr;
  const double* fcr = nullptr;
  const double* fJ = nullptr;
  const double* rule = nullptr;
  const double* cnm = nullptr;
};
// Restricted additive Schwarz blocks
struct Schwarz {
  const int64_t* bptr = nullptr;
  const int32_t* bidx = nullptr;
  const int32_t* bncore = nullptr;
  const int64_t* inv_ptr = nullptr;
  int nblk = 0;
};

// ---- matvec pipeline ----
void fill(hipStream_t s, double* p, int64_t n, double v);
// mode 0: unknowns = x, knowns = 0; 1: knowns (rhs); 2: rigid mode (phi = 1, phi_n = 0)
void scatter(hipStream_t s, int mode, int n, const int32_t* kind, const double* kappa, const double* x,
             const double* phi_k, const double* phin_k, double* phin, double* phi);
// part 0: y = A x over all columns; 1: y = A x over the DOF columns (rounded down to 4); 2: y += A x over the rest
void block_gemv(hipStream_t s, const Gemv& g, const double* mat, const double* x0, const double* x1, double* y0,
                double* y1, int part);
void block_gemv(hipStream_t s, const Gemv& g, const float* mat, const double* x0, const double* x1, double* y0,
                double* y1, int part);
// the FP32 near field folded for A x: g is the FP64 near field's structure (with ncol_dof); writes the folded
// block columns ucol_idx (into [x | L0 | L1], the locals at stride Cl, nbC = boxes x Cl) and weights umat at the
// offsets ucol_ptr / umat_ptr
void fold_near(hipStream_t s, const Gemv& g, int C, int Cl, int n, int64_t nbC, const double* mat, const int32_t* kind,
               const double* kappa, const int64_t* ucol_ptr, const int64_t* umat_ptr, int32_t* ucol_idx, float* umat);
// y = U [x | L0 | L1] over the folded blocks (g: col_ptr / mat_ptr / col_idx of the folded layout); part as above
void block_gemv_u(hipStream_t s, const Gemv& g, const float* mat, int n, int64_t nbC, const double* x,
                  const double* L0, const double* L1, double* y, int part);
// Crows: multipole rows computed (the first Crows coefficients; 0 = all C); fewer than C: the multipoles are
// stored with stride Crows
void p2m(hipStream_t s, int C, int CP, const Gemv& g, const double* mat, const double* x0, const double* x1,
         double* M0, double* M1, int Crows = 0);
void p2m(hipStream_t s, int C, int CP, const Gemv& g, const float* mat, const double* x0, const double* x1,
         double* M0, double* M1, int Crows = 0);
// stage == nullptr: D[dst] += T S[src] (atomics; M2M, L2L); else per-pair results to stage (M2L)
void translate(hipStream_t s, int C, int RP, int KP, const Tiles& t, const double* ops, const double* S0,
               const double* S1, double* D0, double* D1, double* stage);
void translate(hipStream_t s, int C, int RP, int KP, const Tiles& t, const float* ops, const double* S0,
               const double* S1, double* D0, double* D1, float* stage);
// the first C coefficients of every box's two multipoles (stored with stride Cs >= C) as FP32 columns of C
void pack_m32(hipStream_t s, int64_t n, int C, int Cs, const double* M0, const double* M1, float* M32);
// expansion orders the rotation M2L runs the FP32 operator with (m2l_rot_compiled), and also its inner
// operator at a lower order than the plan's (m2l_rot_inner_compiled)
bool m2l_rot_compiled(int P);
bool m2l_rot_inner_compiled(int P);
// rotation M2L of the FP32 operator (throws for an expansion order it is not compiled for)
void m2l_rot(hipStream_t s, int P, const Tiles& t, const float* rot, const float* M32, float* stage);
void m2l_reduce(hipStream_t s, int C, int nboxes, const int32_t* order, const int64_t* tptr, const int32_t* tpairs,
                const double* stage, double* L0, double* L1);
// FP32 staging of C rows per column; the locals are stored with stride Cout >= C, rows [C, Cout) set to zero
void m2l_reduce(hipStream_t s, int C, int Cout, int nboxes, const int32_t* order, const int64_t* tptr,
                const int32_t* tpairs, const float* stage, double* L0, double* L1);
void assemble(hipStream_t s, int mode, int n, const int32_t* kind, const int32_t* point, const double* alpha,
              const double* phi, const double* y0, const double* y1, const double* phi_k, double* out);

// ---- setup ----
void morton_keys(hipStream_t s, int n, const int32_t* point, const double* X, const double bounds[6], uint64_t* keys,
                 int32_t* vals);
// stable LSD radix sort of (key, value) pairs on the low `bits` key bits; tmp == nullptr: only sets tmp_bytes
hipError_t radix_sort_pairs(hipStream_t s, void* tmp, size_t& tmp_bytes, const uint64_t* keys_in, uint64_t* keys_out,
                            const int32_t* vals_in, int32_t* vals_out, int n, int bits);
void p2m_build(hipStream_t s, int P, const P2MBuild& a, double* mat);  // throws for P outside 1..10

// ---- plan tree stage (tree.hip): octree, interaction lists, target boxes ----
// Per level l (boxes [lb, le), faces in per-face arrays of F): tree_split appends one key
// ((box - lb) * 32 + cell) << 32 | face per face whose box splits (more than leaf_max faces and may_split) and
// shifts every face's path key by one digit; after the keys are sorted, tree_heads + an inclusive scan number
// the distinct (box, cell) runs, tree_children creates one child box per run at le + run, and tree_assign
// moves each face into its run's child if the child's inside test holds (else the face stays with its box).
// err |= 1: a box with more than kTreeCells cells, 2 (tree_nonempty): a child that kept no face.
constexpr int kTreeCells = 27;  // cells per box at most (2 per axis; 3 at the root when its side rounds up)
void tree_split(hipStream_t s, int F, int lb, int leaf_max, int may_split, const double* pole, const int32_t* fbox,
                uint64_t* fkey, const BoxGeom* geom, const int32_t* count, uint64_t* out, uint32_t* counter,
                int32_t* err);
void tree_heads(hipStream_t s, int n, const uint64_t* keys, int32_t* flags);
void tree_children(hipStream_t s, int n, const uint64_t* keys, const int32_t* ci, int lb, int le, BoxGeom* geom,
                   int32_t* parent, int32_t* level, uint64_t* pk, int32_t* count, int32_t* nchild,
                   int32_t* first_child, int32_t* child_tab);
void tree_assign(hipStream_t s, int n, const uint64_t* keys, const int32_t* ci, int le, const double* pole,
                 const BoxGeom* geom, int32_t* fbox, uint64_t* fkey, int32_t* count);
void tree_nonempty(hipStream_t s, int nc, int le, const int32_t* count, int32_t* err);
void iota_i32(hipStream_t s, int n, int32_t* out);
void fill_i32(hipStream_t s, int64_t n, int32_t* p, int32_t v);
void fill_u64(hipStream_t s, int64_t n, uint64_t* p, uint64_t v);
// box b's faces are those whose `digits`-digit path key starts with its own path key pk[b] (level[b] digits)
void tree_ranges(hipStream_t s, int nb, int digits, const int32_t* level, const uint64_t* pk, const uint64_t* keys,
                 int F, int64_t* range);
// internal offsets: leaves from their faces, then each level's parents (deepest first) from their children
void tree_leaf_off(hipStream_t s, int nb, const int32_t* nchild, const int64_t* range, const int32_t* perm,
                   const double* maxoff, double* off);
void tree_parent_off(hipStream_t s, int lb, int le, const int32_t* nchild, const int32_t* first_child, double* off);
void tree_targets(hipStream_t s, int N, const double* X, const BoxGeom* geom, const int32_t* nchild,
                  const int32_t* child_tab, int32_t* target);
// near candidates (A, B) of one level, as keys A << 32 | B: list_count flags the near pairs and counts the child
// pairs that are M2L pairs or next level's candidates; list_write writes them at the exclusive scans of the
// counts; list_totals writes the three totals to out[0..2]
void list_count(hipStream_t s, int np, const uint64_t* pairs, const BoxGeom* geom, const double* off,
                const int32_t* nchild, const int32_t* first_child, double theta, int32_t* nearf, int32_t* m2lc,
                int32_t* nextc);
void list_write(hipStream_t s, int np, const uint64_t* pairs, const BoxGeom* geom, const double* off,
                const int32_t* nchild, const int32_t* first_child, double theta, const int32_t* nearf,
                const int32_t* near_at, const int32_t* m2l_at, const int32_t* next_at, uint64_t* near_out,
                uint64_t* m2l_out, uint64_t* next_out);
void list_totals(hipStream_t s, int np, const int32_t* nearf, const int32_t* m2lc, const int32_t* nextc,
                 const int32_t* near_at, const int32_t* m2l_at, const int32_t* next_at, int64_t* out);
// sorted keys A << 32 | B -> CSR over A in [0, nb]: ptr (nb + 1), idx = B
void csr_from_keys(hipStream_t s, int nb, const uint64_t* keys, int64_t n, int64_t* ptr, int32_t* idx);
// ---- near-field blocks on the device (tree.hip), the plan's near structure from the tree stage ----
// target nodes grouped by box (node_keys, then a sort and csr_from_keys); boxes with targets (flags, scan,
// compact); per target box its ranges into the face order (own box, near boxes, near leaves of ancestors:
// count, scan, write, sort, merge in place; err |= 4 on overlapping ranges), then its DOF columns (keys per face
// corner, sorted, unique, CSR); chunks of <= kNearRows rows in descending cost (err |= 8 if a chunk's cost
// reaches 2^22); their sizes scanned into the block offsets; fill; the L2P coefficients per row.
constexpr int kNearRows = 32;
void node_keys(hipStream_t s, int N, const int32_t* target, uint64_t* keys);
void ns_tbox_flags(hipStream_t s, int nb, const int64_t* tptr, int32_t* flags);
void ns_tbox_compact(hipStream_t s, int nb, const int32_t* flags, const int32_t* at, int32_t* tbox);
void ns_chunk_count(hipStream_t s, int ntb, const int64_t* tptr, const int32_t* tbox, int32_t* cnt);
// (ns_range_count, ns_range_merge and ns_chunk_count also write entry ntb: count 0, for the scans' totals)
void ns_range_count(hipStream_t s, int ntb, const int32_t* tbox, const int32_t* parent, const int64_t* near_ptr,
                    const int32_t* near_idx, const int32_t* nchild, int32_t* cnt);
void ns_range_write(hipStream_t s, int ntb, const int32_t* tbox, const int32_t* parent, const int64_t* near_ptr,
                    const int32_t* near_idx, const int32_t* nchild, const int64_t* box_range, const int32_t* at,
                    uint64_t* keys, int32_t* ends);
void ns_range_merge(hipStream_t s, int ntb, const int32_t* at, const int32_t* cnt, const uint64_t* keys,
                    const int32_t* ends, int64_t* mstart, int64_t* mend, int32_t* mcnt, int64_t* nfaces, int32_t* err);
void ns_dof_keys(hipStream_t s, int npf, int64_t total, int ntb, const int64_t* fstart, const int32_t* at,
                 const int32_t* mcnt, const int64_t* mstart, const int64_t* mend, const int32_t* fdof, uint64_t* keys);
void unique_flags(hipStream_t s, int64_t n, const uint64_t* keys, int32_t* flags);
void compact_u64(hipStream_t s, int64_t n, const uint64_t* keys, const int32_t* flags, const int32_t* at,
                 uint64_t* out);
void ns_chunk_keys(hipStream_t s, int ntb, const int64_t* tptr, const int32_t* tbox, const int32_t* chunk_at,
                   const int64_t* dptr, int C, uint64_t* keys, int32_t* err);
void ns_chunk_ti(hipStream_t s, int ntb, const int32_t* chunk_at, int32_t* chunk_ti);
void ns_chunk_sizes(hipStream_t s, int nch, const uint64_t* order, const int32_t* chunk_ti, const int32_t* chunk_at,
                    const int64_t* tptr, const int32_t* tbox, const int64_t* dptr, const int32_t* mcnt, int C,
                    int64_t* rows_v, int64_t* cols_v, int64_t* mat_v, int64_t* rng_v, int64_t* nnz_v);
void ns_fill(hipStream_t s, int nch, const uint64_t* order, const int32_t* chunk_ti, const int32_t* chunk_at,
             const int64_t* tptr, const int32_t* tnodes, const int32_t* tbox, const int64_t* dptr, const int32_t* dofs,
             const int32_t* rat, const int32_t* mcnt, const int64_t* mstart, const int64_t* mend, int C, int n_dof,
             const int64_t* row_ptr, const int64_t* col_ptr, const int64_t* range_ptr, int32_t* row_out,
             int32_t* row_block, int64_t* node_row, int32_t* col_idx, int32_t* ncol_dof, int64_t* ranges,
             int64_t* range_prefix, const int64_t* fstart, int64_t* block_fstart);
// local DOF column of each face node of the target boxes' merged ranges (npf per face, from npf fstart[ti]; -1 for
// an edge midpoint without a DOF)
void ns_face_cols(hipStream_t s, int npf, int64_t total, int ntb, const int64_t* fstart, const int32_t* at,
                  const int32_t* mcnt, const int64_t* mstart, const int64_t* mend, const int32_t* fdof,
                  const int64_t* dptr, const int32_t* dofs, int32_t* cols);
void l2p_coef(hipStream_t s, int64_t nrows, int P, const int32_t* row_out, const int32_t* target, const double* X,
              const BoxGeom* geom, const int32_t* level, const double* level_side, const double* cnm, double* out);
// rocPRIM radix sorts and scans; tmp == nullptr: only sets tmp_bytes
hipError_t scan_exclusive_i64(hipStream_t s, void* tmp, size_t& tmp_bytes, const int64_t* in, int64_t* out,
                              int64_t n);
hipError_t sort_keys_u64(hipStream_t s, void* tmp, size_t& tmp_bytes, const uint64_t* in, uint64_t* out, int64_t n,
                         int begin_bit, int end_bit);
hipError_t sort_pairs_u64_i32(hipStream_t s, void* tmp, size_t& tmp_bytes, const uint64_t* keys_in,
                              uint64_t* keys_out, const int32_t* vals_in, int32_t* vals_out, int64_t n, int begin_bit,
                              int end_bit);
hipError_t scan_inclusive_i32(hipStream_t s, void* tmp, size_t& tmp_bytes, const int32_t* in, int32_t* out,
                              int64_t n);
hipError_t scan_exclusive_i32(hipStream_t s, void* tmp, size_t& tmp_bytes, const int32_t* in, int32_t* out,
                              int64_t n);
void near_quad(hipStream_t s, int64_t nrows, const NearQuad& a, const Gemv& g, double* mat);
void to_f32(hipStream_t s, int64_t n, const double* a, float* b);
void diag_inv(hipStream_t s, int n, const NearRows& a, double* dinv);
void schwarz_build(hipStream_t s, const Schwarz& b, int max_block, const NearRows& a, double* inv, int32_t* failed);

// ---- GMRES vectors and preconditioner apply ----
// partial[i * kDotBlocks + blk] = sum over the block's chunk of V_i . w, i < k
void dots(hipStream_t s, int k, const double* V, int64_t ldv, const double* w, int64_t n, double* partial);
void reduce_partials(hipStream_t s, const double* partial, int k, double* out);
void vsum(hipStream_t s, const double* V, int64_t ldv, int k, const double* h, double sign, double* w, int64_t n);
// out[0..nk) = P in[0..nk) (P: nk x nk, row-major), out[nk..nb) = in[nk..nb); in and out must not overlap
void proj_coef(hipStream_t s, int nk, int nb, const double* P, const double* in, double* out);
void axpby(hipStream_t s, int64_t n, double a, const double* x, double b, double* y);
void sub(hipStream_t s, int64_t n, const double* a, const double* b, double* out);
void mul(hipStream_t s, int64_t n, const double* d, double* y);
void mul_to(hipStream_t s, int64_t n, const double* d, const double* x, double* y);
void scale_dup(hipStream_t s, int64_t n, double a, double* w, double* dup);
void schwarz_apply(hipStream_t s, const Schwarz& b, const double* inv, const double* r, double* z);

// ILU(0) triangular sweeps (ILU0_CRS::solve, basic_linear_systems.hpp:2562-2603) on the factored rows, level by
// level. One sweep segment is either one level with many rows (one thread per row over the grid) or a run of
// consecutive small levels done by one workgroup with a barrier between levels. z holds the right-hand side on
// entry to the forward sweep and the solution after the backward sweep.
struct IluRows {
  const int64_t* ptr = nullptr;   // n + 1
  const int32_t* col = nullptr;
  const double* val = nullptr;
  const int64_t* diag = nullptr;  // position of each row's diagonal
};
constexpr int kIluBigLevel = 4096;  // levels with at least this many rows get a launch of their own
// Schwarz blocks from the rows of the reference's sparse matrix (its SCHWARZ option): rows_out[b] rows of
// A_b^-1 stored per block (all m for the averaged combination)
void schwarz_build_rows(hipStream_t s, const Schwarz& b, int max_block, const IluRows& A, const int32_t* rows_out,
                        double* inv, int32_t* failed);
// averaged additive Schwarz: y (sum of block sizes) receives every block's local solution, z[g] the mean of
// the entries of g (occurrence list occ_pos[occ_ptr[g] .. occ_ptr[g+1]) into y, in block order)
void schwarz_apply_averaged(hipStream_t s, const Schwarz& b, int n, const int64_t* occ_ptr, const int64_t* occ_pos,
                            const double* inv, const double* r, double* y, double* z);
constexpr int kIluRunThreads = 1024;
// backward: upper part and the division by the diagonal, else the strictly lower part.
// One level: its nrows rows at `level_rows`, one thread each.
void ilu_level(hipStream_t s, const IluRows& A, const int32_t* level_rows, int64_t nrows, bool backward, double* z);
// Levels [l0, l1) in one workgroup (rows[lptr[l] .. lptr[l+1]) per level).
void ilu_run(hipStream_t s, const IluRows& A, const int32_t* rows, const int32_t* lptr, int l0, int l1, bool backward,
             double* z);
void perm_gather(hipStream_t s, int n, const int32_t* perm, const double* in, double* out);

// ---- allocation-free work memory ----
// The file-scope device array `scratch` of kernels.hip (kScratchBytes, 256-byte aligned), one copy per GPU,
// held until the process ends: *ptr = its device address on the current device (looked up once per device
// with hipGetSymbolAddress and kept), *bytes = its size. Every caller on a GPU shares that GPU's array.
constexpr size_t kScratchBytes = size_t(64) << 20;
hipError_t scratch_region(void** ptr, size_t* bytes);
void perm_scatter(hipStream_t s, int n, const int32_t* perm, const double* in, double* out);

}  // namespace dev
}  // namespace bemfmm
