// Host side of the MI355X-native FMM-accelerated GMRES BVP engine and its C ABI (include/bemfmm.h):
// contexts, device buffers, the per-matvec launch sequence (captured into HIP graphs), setup, GMRES and
// iterative refinement. Device code lives in kernels.hip behind the launchers of kernels.hpp; this file
// uses only the HIP runtime API.
//
// Matvec launch sequence on one HIP stream, all data resident in HBM (the reference's compute_Ax_minus_b,
// bem/BEM_solveBVP_GMRES_FMM.hpp:34-158, with updateFMM, lib/include/moments.hpp:1547-1574): scatter,
// P2M, M2M, M2L (+ per-box reduction), L2L, near field + L2P (the near-field DOF columns on a second
// stream beside the FMM passes), assemble.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <future>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/bemfmm.h"
#include "kernels.hpp"
#include "plan.hpp"

namespace {

using bemfmm::BoxGeom;
using bemfmm::Plan;
namespace dev = bemfmm::dev;
using dev::kDotBlocks;
using dev::kRotTP;
using dev::kSchwarzMaxBlock;
using dev::kTP;
using dev::kTPM2L;

struct HipError : std::runtime_error {
  int code;
  HipError(const std::string& s, int c) : std::runtime_error(s), code(c) {}
};

#define HIPCHECK(expr)                                                                                        \
  do {                                                                                                        \
    hipError_t e_ = (expr);                                                                                   \
    if (e_ != hipSuccess)                                                                                     \
      throw HipError(std::string(#expr) + ": " + hipGetErrorString(e_),                                      \
                     e_ == hipErrorOutOfMemory ? BEMFMM_E_OOM : BEMFMM_E_DEVICE);                             \
  } while (0)

double now() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

// Stream capture against device-wide synchronisation, process-wide. hipFree and hipHostFree perform an
// implicit hipDeviceSynchronize (HIP runtime API documentation), which synchronises every stream of the
// device, including a stream that another thread is capturing into a graph at that moment: a
// synchronisation that stream capture does not allow. With several contexts driven from several host
// threads (the request loop), one context's frees (re-prepare, Krylov basis growth, teardown) would
// otherwise run while another context captures its matvec graph. Captures (begin, end, instantiate) hold
// the gate exclusively; every allocation and free of this library holds it shared.
std::shared_mutex g_capture_gate;

hipError_t dev_malloc(void** p, size_t bytes) {
  std::shared_lock<std::shared_mutex> lk(g_capture_gate);
  return hipMalloc(p, bytes);
}
hipError_t dev_free(void* p) {
  std::shared_lock<std::shared_mutex> lk(g_capture_gate);
  return hipFree(p);
}
hipError_t pinned_malloc(void** p, size_t bytes) {
  std::shared_lock<std::shared_mutex> lk(g_capture_gate);
  return hipHostMalloc(p, bytes, hipHostMallocDefault);
}
hipError_t pinned_free(void* p) {
  std::shared_lock<std::shared_mutex> lk(g_capture_gate);
  return hipHostFree(p);
}

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------
template <typename T>
// Device array of n elements in an allocation of cap >= n elements. alloc() keeps the allocation when it is
// large enough: a re-prepare after nodes moved (every RK stage of a session) needs about the same sizes, so
// it reallocates nothing large; a larger request allocates 1/8 more than asked, so that the next stages'
// small growth fits. Besides the frees and allocations saved, this matters on the ROCm 7 runtime: buffers
// that a context's matvec graphs had used stayed allocated for a while after they were freed (4.1 GB per
// re-prepare of an 80k-node tank, measured on the MI355X).
struct DBuf {
  T* p = nullptr;
  size_t n = 0, cap = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), n(o.n), cap(o.cap) {
    o.p = nullptr;
    o.n = o.cap = 0;
  }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      cap = o.cap;
      o.p = nullptr;
      o.n = o.cap = 0;
    }
    return *this;
  }
  void swap(DBuf& o) noexcept {
    std::swap(p, o.p);
    std::swap(n, o.n);
    std::swap(cap, o.cap);
  }
  ~DBuf() { release(); }
  void release() {
    if (p) {
      const hipError_t e = dev_free(p);
      if (e != hipSuccess)  // the memory stays allocated: say so instead of losing it silently
        std::fprintf(stderr, "[bemfmm] hipFree of %zu bytes failed: %s\n", n * sizeof(T), hipGetErrorString(e));
    }
    p = nullptr;
    n = cap = 0;
  }
  void alloc(size_t count) {
    if (count == 0) count = 1;
    if (p && count <= cap) {
      n = count;
      return;
    }
    const size_t want = p ? count + count / 8 : count;
    release();
    HIPCHECK(dev_malloc((void**)&p, want * sizeof(T)));
    n = count;
    cap = want;
  }
  // Grows to at least `count` elements keeping the first `keep` (device-to-device copy on s, completed before
  // the old allocation is freed); a larger allocation takes half as much again for the next growth.
  void reserve_keep(size_t count, size_t keep, hipStream_t s) {
    if (p && count <= cap) {
      n = std::max(n, count);
      return;
    }
    const size_t want = std::max<size_t>(count + count / 2, 1);
    T* q = nullptr;
    HIPCHECK(dev_malloc((void**)&q, want * sizeof(T)));
    if (p && keep) {
      HIPCHECK(hipMemcpyAsync(q, p, std::min(keep, n) * sizeof(T), hipMemcpyDeviceToDevice, s));
      HIPCHECK(hipStreamSynchronize(s));
    }
    release();
    p = q;
    n = count;
    cap = want;
  }
  // Setup-time download of the first `count` elements; completes before returning.
  void download(std::vector<T>& v, size_t count, hipStream_t s) const {
    v.resize(count);
    if (count) {
      HIPCHECK(hipMemcpyAsync(v.data(), p, count * sizeof(T), hipMemcpyDeviceToHost, s));
      HIPCHECK(hipStreamSynchronize(s));
    }
  }
  // Setup-time upload; completes before returning (the source may be a temporary).
  void upload(const std::vector<T>& v, hipStream_t s) {
    alloc(v.size());
    if (!v.empty()) {
      HIPCHECK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
      HIPCHECK(hipStreamSynchronize(s));
    }
  }
  size_t bytes() const { return cap * sizeof(T); }
};

struct DGemv {
  DBuf<int64_t> row_ptr, col_ptr, mat_ptr;
  DBuf<int32_t> row_out, col_idx;
  DBuf<double> mat;
  int nblk = 0;
  void upload(const bemfmm::BlockGemv& g, hipStream_t s) {
    row_ptr.upload(g.row_ptr, s);
    col_ptr.upload(g.col_ptr, s);
    mat_ptr.upload(g.mat_ptr, s);
    row_out.upload(g.row_out, s);
    col_idx.upload(g.col_idx, s);
    nblk = g.n_blocks();
  }
  size_t bytes() const {
    return row_ptr.bytes() + col_ptr.bytes() + mat_ptr.bytes() + row_out.bytes() + col_idx.bytes() + mat.bytes();
  }
  dev::Gemv view(const int32_t* ncol_dof = nullptr) const {
    dev::Gemv g;
    g.row_ptr = row_ptr.p;
    g.col_ptr = col_ptr.p;
    g.mat_ptr = mat_ptr.p;
    g.row_out = row_out.p;
    g.col_idx = col_idx.p;
    g.ncol_dof = ncol_dof;
    g.nblk = nblk;
    return g;
  }
};

struct DTiles {
  DBuf<int32_t> op, dst, src;
  DBuf<int64_t> p0, p1;
  int ntiles = 0;
  int tp = kTP;  // pairs per tile
  // slot[op] = index of operator op in the context's device operator table
  void upload(const bemfmm::PairGroups& g, hipStream_t s, const std::vector<int32_t>& slot, int pairs_per_tile = kTP) {
    tp = pairs_per_tile;
    std::vector<int32_t> top;
    std::vector<int64_t> tp0, tp1;
    for (size_t k = 0; k < g.op.size(); ++k)
      for (int64_t a = g.pair_ptr[k]; a < g.pair_ptr[k + 1]; a += tp) {
        top.push_back(slot[g.op[k]]);
        tp0.push_back(a);
        tp1.push_back(std::min(g.pair_ptr[k + 1], a + (int64_t)tp));
      }
    ntiles = (int)top.size();
    op.upload(top, s);
    p0.upload(tp0, s);
    p1.upload(tp1, s);
    dst.upload(g.dst, s);
    src.upload(g.src, s);
  }
  size_t bytes() const { return op.bytes() + dst.bytes() + src.bytes() + p0.bytes() + p1.bytes(); }
  dev::Tiles view() const {
    dev::Tiles t;
    t.op = op.p;
    t.p0 = p0.p;
    t.p1 = p1.p;
    t.dst = dst.p;
    t.src = src.p;
    t.ntiles = ntiles;
    t.tp = tp;
    return t;
  }
};

}  // namespace

struct bemfmm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bemfmm_opts opts{};
  std::string err;
  // host inputs
  int n_nodes = 0, n_faces = 0, n_dof = 0;
  std::vector<double> X;
  std::vector<int32_t> tri, face_dof, dof_kind, dof_point;
  std::vector<double> phi_bc, phin_bc;
  std::vector<double> robin_bc;  // kappa per DOF (caller order, physical units); empty = 0
  bool have_geom = false, have_bc = false, prepared = false;
  // internal DOF numbering (Morton order unless opts.dof_reorder == 0): internal i is caller DOF perm[i]
  std::vector<int32_t> perm, inv, face_dof_i, kind_i, point_i;
  bool perm_identity = true;
  DBuf<int32_t> d_perm;
  // plan and device state
  std::unique_ptr<Plan> plan;
  int C = 0, n_boxes = 0, RP = 0, KP = 0, p2m_rows_pad = 0;
  DGemv near, p2m;
  std::vector<DTiles> m2m, l2l;
  DTiles m2l;
  DBuf<int64_t> m2l_tptr;    // per box: its M2L pairs in m2l_tpairs
  DBuf<int32_t> m2l_rorder;  // boxes in descending M2L pair count (launch order of the reduction)
  DBuf<int32_t> m2l_tpairs;
  DBuf<double> m2l_stage;    // per-pair M2L results [pair][layer][C]
  DBuf<double> ops;
  // FP32 copies for BEMFMM_PRECISION_FP32_REFINE (inner matvecs); the M2L staging is shared
  DBuf<float> near_mat32, p2m_mat32, ops32;  // near_mat32: the FP32 near field folded for A x (k_fold_near)
  DBuf<int64_t> nu_col_ptr, nu_mat_ptr;     // its blocks' column and weight offsets
  DBuf<int32_t> nu_col_idx;                 // its columns, into [x | L0 | L1]
  // device operator table, [slot][KP][RP], kept across prepares: slot of each (kind, a, b, c) key
  std::map<std::array<int, 4>, int32_t> op_slot;
  int ops_cap = 0, ops_used = 0, ops_P = -1, ops32_valid = 0;
  int ops_uploaded = 0;  // operators uploaded by the last prepare
  std::vector<int32_t> m2l_slot;  // plan op -> table slot (for re-tiling by the autotuner)
  int m2l_tp = 0;                // pairs per M2L tile in use
  int gemv_var[2] = {0, 0};       // near GEMV launch shape (dev::Gemv::variant) for the FP64 / FP32 operator
  int p2m_var[2] = {0, 0};        // P2M launch shape for the FP64 / FP32 operator
  // rotation M2L of the FP32 operator: factors [slot][RS] (slots of the operator table; M2L slots only)
  bool m2l_rot = false;
  // FP32 rotation factors [operator slot][RS] of the rotation M2L at the plan's order (rot) and at the inner
  // operator's lower order P_in (rot_in; P_in = 0: the inner operator uses the plan's order)
  struct RotTable {
    DBuf<float> f;
    std::vector<char> have;
    int RS = 0;
  } rot, rot_in;
  int P_in = 0;
  // the FP32 operator at the inner order runs its expansions compact: (P_in+1)^2 coefficients per box
  // (C_in, its padded RP_in / KP_in), M2M and L2L on the leading blocks of the 16 octant operators (ops_in):
  // the tiles of m2m / l2l with their own operator index per tile (m2m_in_op / l2l_in_op index ops_in)
  int C_in = 0, RP_in = 0, KP_in = 0;
  DBuf<double> ops_in;
  std::vector<DBuf<int32_t>> m2m_in_op, l2l_in_op;
  DTiles m2l_rt;  // M2L tiles of kRotTP pairs (one offset each)
  bool m2l_pw_on = false;  // FP32 operator's M2L through the plane-wave composed operators (m2l_kernel 3)
  DTiles m2l_pw;           // its tiles (operator slots of the plane-wave M2L offsets)
  int m2l_min_level = 0;  // coarsest level with M2L pairs: M2M into and L2L out of coarser levels are skipped
  DBuf<float> m32;  // FP32 multipoles [box][layer][C] read by the rotation M2L
  // near field beside the FMM passes: the near-only GEMV runs on stream2 between a fork and a join event
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // The matvec graphs are captured and replayed on streams of their own (gstream, its fork gstream2; events
  // gfork/gjoin for the capture, gin/gout to order the replays with the context's stream), so that no
  // kernel launched eagerly shares a stream with a capture or a replay: with eager kernels on those streams
  // the ROCm 7 runtime kept buffers used by the graphs allocated after they were freed.
  hipStream_t gstream = nullptr, gstream2 = nullptr;
  hipEvent_t ev_gfork = nullptr, ev_gjoin = nullptr, ev_gin = nullptr, ev_gout = nullptr;
  hipEvent_t prep_ev[2] = {};  // around the near-field quadrature of a prepare (timing)
  DBuf<int32_t> near_ncd;  // DOF columns per near block
  double* op_stage = nullptr;  // pinned host staging for operator uploads
  size_t op_stage_n = 0;
  double* hres = nullptr;      // pinned host copy of hbuf (the Arnoldi column of one iteration)
  DBuf<int32_t> kind, point;
  DBuf<double> alpha, phi_k, phin_k;
  DBuf<double> robin;  // kappa L per DOF (internal order; 0 except on ROBIN DOFs)
  DBuf<double> buf0, buf1;   // [phi_n | L0], [phi | L1]
  DBuf<double> M0, M1;
  DBuf<double> y0, y1;
  DBuf<double> xin, yout, bvec, tmpv, xsol, xbase, jacobi, rvec, dvec, ptmp;
  DBuf<double> gin, gout;            // fixed input/output of the captured matvec graphs
  const double* gin_holds = nullptr;  // a vector gin already equals (valid until the next evaluate)
  hipGraphExec_t gexec[2] = {};      // [FP64, FP32] matvec graphs
  // restricted additive Schwarz: blocks and the core rows of the block inverses
  bemfmm::SchwarzBlocks sw_host;
  // the mesh-graph Schwarz blocks (schwarz_variant 0) in the caller's DOF numbering, with the key of the topology
  // and options they were built for: a re-prepare after nodes moved keeps them (they depend on the DOF graph only)
  bemfmm::SchwarzBlocks sw_caller;
  uint64_t sw_topo_key = 0;
  int sw_failed = 0;
  DBuf<int64_t> sw_ptr, sw_invptr;
  DBuf<int32_t> sw_idx, sw_ncore;
  DBuf<double> sw_inv;
  // the reference's SCHWARZ (schwarz_variant 1): every block's full inverse, the DOFs' occurrences in the blocks
  DBuf<int64_t> sw_occ_ptr, sw_occ_pos;
  DBuf<int32_t> sw_rows_out;
  DBuf<double> sw_y;
  // ILU(0) of the reference's sparse preconditioner matrix: factored rows and the sweeps' level schedules
  struct IluSegment {
    int l0, l1;    // levels [l0, l1) of the sweep
    bool single;   // one large level: one thread per row over the grid
  };
  DBuf<int64_t> ilu_ptr, ilu_diag;
  DBuf<int32_t> ilu_col, ilu_lrows, ilu_lptr, ilu_urows, ilu_uptr;
  DBuf<double> ilu_val;
  std::vector<int32_t> ilu_lptr_h, ilu_uptr_h;
  std::vector<IluSegment> ilu_lseg, ilu_useg;
  int64_t ilu_nnz = 0;
  int ilu_levels = 0;
  // setup inputs kept with the context (grow-only, reused by every re-prepare: a free would synchronise the
  // whole device, including the other contexts' work)
  struct SetupBufs {
    DBuf<double> dX;
    DBuf<int32_t> dpt, v0, v1;
    DBuf<uint64_t> k0, k1;
    DBuf<unsigned char> sort_tmp;
    DBuf<int32_t> row_block, ncol_dof, face_perm, fv, fdof;
    DBuf<int64_t> range_ptr, ranges, range_prefix;
    DBuf<double> Xs, fX, fcent, fcr, fJ, l2p, d_dun, d_dun6;
    DBuf<int64_t> fptr;
    DBuf<int32_t> pfaces, plc;
    DBuf<double> pcs, dcnm;
    DBuf<int64_t> node_row;
    DBuf<int32_t> failed;
    DTiles alt_tiles;  // the autotuner's alternative M2L tiling
  } sb;
  // the plan's tree stage on the device (plan_tree_lists_device)
  struct TreeBufs {
    DBuf<double> pole, maxoff, X, off;
    DBuf<BoxGeom> geom;
    DBuf<int32_t> fbox, flags, ci, vals, perm, parent, level, count, nchild, first_child, child_tab, target;
    DBuf<int32_t> nearf, m2lc, nextc, near_at, m2l_at, next_at, near_idx, m2l_idx, err;
    DBuf<uint64_t> fkey, keys0, keys1, pk, pairs, next, near_all, m2l_all;
    DBuf<int64_t> range, near_ptr, m2l_ptr, tot;
    DBuf<uint32_t> ctr;
  } tb;
  // the near-field structure built on the device (near_struct_device)
  struct NearBufs {
    DBuf<uint64_t> k0, k1, rkeys, rkeys2, dk0, dk1, ck0, ck1;
    DBuf<int64_t> tptr, mstart, mend, nfaces, fstart, dptr, v_rows, v_cols, v_mat, v_rng, v_nnz, nnzp;
    DBuf<int32_t> tnodes, flags, at, tbox, rcnt, rat, rends, rends2, mcnt, uflag, uat, dofs, ccnt, cat, cti;
    DBuf<int32_t> face_cols;       // local column of each face corner per target box (near quadrature)
    DBuf<int64_t> block_fstart;    // per near block: its target box's first entry in face_cols / 3
    DBuf<double> lside, cnm;
  } nbuf;
  int64_t near_rows = 0;  // rows of the near-field blocks
  DBuf<double> V;            // Krylov basis
  int64_t v_cap = 0;
  int basis_n = -1;          // DOF count the basis is laid out for
  DBuf<double> partial, hbuf;
  DBuf<double> gcro_proj;  // Q_C Q_C^T over the kept basis vectors of a recycled inner solve (nk x nk)
  bemfmm_stats stats{};
  // profiling
  bool prof = false;
  hipEvent_t ev[9] = {};
  int n_matvec = 0;
  std::vector<double> history;         // GMRES residual estimate per iteration (last solve)
  std::vector<double> refine_history;  // FP64 residual at each refinement step (last solve, FP32_REFINE)
};

namespace {

void set_err(bemfmm_ctx* c, const std::string& s) {
  if (c) c->err = s;
}

template <typename F>
int guarded(bemfmm_ctx* c, F&& f) {
  try {
    if (c && c->device >= 0) HIPCHECK(hipSetDevice(c->device));
    return f();
  } catch (const HipError& e) {
    set_err(c, e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_err(c, "host allocation failed");
    return BEMFMM_E_OOM;
  } catch (const std::exception& e) {
    set_err(c, e.what());
    return BEMFMM_E_INVALID;
  }
}

dev::Schwarz schwarz_view(const bemfmm_ctx* c) {
  dev::Schwarz b;
  b.bptr = c->sw_ptr.p;
  b.bidx = c->sw_idx.p;
  b.bncore = c->sw_ncore.p;
  b.inv_ptr = c->sw_invptr.p;
  b.nblk = c->sw_host.n_blocks();
  return b;
}

template <typename T>
void launch_gemv(bemfmm_ctx* c, DGemv& g, const T* mat, const double* x0, const double* x1, double* y0, double* y1,
                 int part = 0, hipStream_t s = nullptr) {
  dev::Gemv v = g.view(c->near_ncd.p);
  v.variant = c->gemv_var[std::is_same<T, float>::value ? 1 : 0];
  dev::block_gemv(s ? s : c->stream, v, mat, x0, x1, y0, y1, part);
}

// coefficients per box of the FP32 operator's expansions (the inner order's when one is set)
int fp32_coeffs(const bemfmm_ctx* c) { return (c->m2l_rot && c->P_in) ? c->C_in : c->C; }

// y = U [x | L0 | L1] over the folded FP32 near field (A x of the FP32 operator: y holds y0 - y1)
void launch_gemv_u(bemfmm_ctx* c, const double* x, const double* L0, const double* L1, double* y, int part,
                   hipStream_t s) {
  dev::Gemv v = c->near.view(c->near_ncd.p);
  v.col_ptr = c->nu_col_ptr.p;
  v.mat_ptr = c->nu_mat_ptr.p;
  v.col_idx = c->nu_col_idx.p;
  v.variant = c->gemv_var[1];
  dev::block_gemv_u(s, v, c->near_mat32.p, c->n_dof, (int64_t)c->n_boxes * fp32_coeffs(c), x, L0, L1, y, part);
}

template <typename T>
void launch_p2m(bemfmm_ctx* c, const T* mat, const double* x0, const double* x1, double* M0, double* M1,
                hipStream_t s) {
  dev::Gemv v = c->p2m.view();
  const bool f32 = std::is_same<T, float>::value;
  v.variant = c->p2m_var[f32 ? 1 : 0];
  // the FP32 operator at an inner order computes the multipoles' first C_in coefficients, stored at stride C_in
  const int rows = (f32 && c->m2l_rot && c->P_in) ? c->C_in : 0;
  dev::p2m(s, c->C, c->p2m_rows_pad, v, mat, x0, x1, M0, M1, rows);
}

template <typename T>
void launch_translate(bemfmm_ctx* c, DTiles& t, const T* ops, const double* S0, const double* S1, double* D0, double* D1,
                      T* stage = nullptr, hipStream_t s = nullptr) {
  dev::translate(s ? s : c->stream, c->C, c->RP, c->KP, t.view(), ops, S0, S1, D0, D1, stage);
}

// the FP32 operator's M2L by rotation, at the inner order P_in when one is set: the first C_in coefficients of
// every multipole (degrees <= P_in; the packing is degree-major) are rotated, translated and rotated back
void launch_m2l_rot(bemfmm_ctx* c, hipStream_t s, float* stage, const double* M0, const double* M1, float* m32) {
  DTiles& t = c->m2l_rt;
  if (t.ntiles == 0) return;
  const int P = c->P_in ? c->P_in : c->plan->P;
  const bemfmm_ctx::RotTable& R = c->P_in ? c->rot_in : c->rot;
  if (R.RS != dev::rot_size(P)) throw std::runtime_error("rotation M2L: factor table stride mismatch");
  const int Cm = (P + 1) * (P + 1);
  dev::pack_m32(s, (int64_t)c->n_boxes * 2 * Cm, Cm, fp32_coeffs(c), M0, M1, m32);
  dev::m2l_rot(s, P, t.view(), R.f.p, m32, stage);
}

void fill(bemfmm_ctx* c, double* p, int64_t n, double v, hipStream_t s = nullptr) {
  dev::fill(s ? s : c->stream, p, n, v);
}

// One evaluation of the BIE operator. mode: 0 = A x (x = in), 1 = rhs, 2 = rigid-mode alpha.
// lo: FP32 near field, P2M and translations (FP32 MFMA), FP64 vectors and accumulation across boxes.
// Launches only (no host synchronisation), so that it can be captured into a HIP graph.
// Work buffers of one evaluation: [phi_n | L0] and [phi | L1] (the DOF values followed by the box locals,
// so that L2P is a column gather like the near field), the multipoles, the per-node near-field sums, the
// per-pair M2L staging (nullptr: the FP64 M2L accumulates into the locals with atomics instead) and the
// FP32 multipoles of the rotation M2L; the stream the evaluation is launched on, the stream of its near-field
// fork and the fork/join events.
struct MatvecWork {
  double *buf0 = nullptr, *buf1 = nullptr, *M0 = nullptr, *M1 = nullptr, *y0 = nullptr, *y1 = nullptr;
  void* stage = nullptr;
  float* m32 = nullptr;
  hipStream_t stream = nullptr, stream2 = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};
MatvecWork context_work(bemfmm_ctx* c) {
  MatvecWork w;
  w.stream = c->stream;
  w.stream2 = c->stream2;
  w.fork = c->ev_fork;
  w.join = c->ev_join;
  w.buf0 = c->buf0.p;
  w.buf1 = c->buf1.p;
  w.M0 = c->M0.p;
  w.M1 = c->M1.p;
  w.y0 = c->y0.p;
  w.y1 = c->y1.p;
  w.stage = c->m2l_stage.p;
  w.m32 = c->m32.p;
  return w;
}

void evaluate_launch(bemfmm_ctx* c, int mode, const double* in, double* out, bool lo, const MatvecWork& w) {
  Plan& pl = *c->plan;
  const int n = c->n_dof;
  const int64_t nbC = (int64_t)c->n_boxes * c->C;
  if (!w.stage && lo) throw std::runtime_error("internal: the FP32 operator needs the M2L staging");
  if (lo && mode != 0) throw std::runtime_error("internal: the FP32 operator evaluates A x only");
  hipStream_t st = w.stream;
  double* phin = w.buf0;
  double* phi = w.buf1;
  double* L0 = w.buf0 + n;
  double* L1 = w.buf1 + n;
  auto rec = [&](int i) {
    if (c->prof) HIPCHECK(hipEventRecord(c->ev[i], st));
  };
  rec(0);
  dev::scatter(st, mode == 0 ? 0 : (mode == 1 ? 1 : 2), n, c->kind.p, c->robin.p, in, c->phi_k.p, c->phin_k.p,
               phin, phi);
  // the near-field columns need only the scattered DOF values: they run on stream2 beside the FMM passes
  // (not while profiling, which times the stages one after the other)
  const bool split = c->opts.overlap_near && !c->prof && c->near.nblk > 0;
  if (split) {
    HIPCHECK(hipEventRecord(w.fork, st));
    HIPCHECK(hipStreamWaitEvent(w.stream2, w.fork, 0));
    if (lo) launch_gemv_u(c, in, L0, L1, w.y0, 1, w.stream2);
    else launch_gemv(c, c->near, c->near.mat.p, w.buf0, w.buf1, w.y0, w.y1, 1, w.stream2);
    HIPCHECK(hipEventRecord(w.join, w.stream2));
  }
  // the FP32 operator at the inner order: expansions of C_in coefficients per box throughout (compact)
  const bool compact = lo && c->m2l_rot && c->P_in;
  const int64_t nbCw = compact ? (int64_t)c->n_boxes * c->C_in : nbC;
  fill(c, w.M0, nbCw, 0., st);
  fill(c, w.M1, nbCw, 0., st);
  rec(1);
  if (lo) launch_p2m(c, c->p2m_mat32.p, phin, phi, w.M0, w.M1, st);
  else launch_p2m(c, c->p2m.mat.p, phin, phi, w.M0, w.M1, st);
  rec(2);
  // the compact translations use the tiles of m2m / l2l with their ops_in operator indices
  auto compact_view = [&](const DTiles& t, const DBuf<int32_t>& op) {
    dev::Tiles v = t.view();
    v.op = op.p;
    return v;
  };
  auto translate_levels = [&](bool up, double* A0, double* A1) {
    if (up) {
      for (int l = pl.n_levels - 1; l >= 1 && l - 1 >= c->m2l_min_level; --l) {
        if (compact)
          dev::translate(st, c->C_in, c->RP_in, c->KP_in, compact_view(c->m2m[l], c->m2m_in_op[l]), c->ops_in.p, A0,
                         A1, A0, A1, (double*)nullptr);
        else
          launch_translate(c, c->m2m[l], c->ops.p, A0, A1, A0, A1, (double*)nullptr, st);
      }
    } else {
      for (int l = std::max(1, c->m2l_min_level + 1); l < pl.n_levels; ++l) {
        if (compact)
          dev::translate(st, c->C_in, c->RP_in, c->KP_in, compact_view(c->l2l[l], c->l2l_in_op[l]), c->ops_in.p, A0,
                         A1, A0, A1, (double*)nullptr);
        else
          launch_translate(c, c->l2l[l], c->ops.p, A0, A1, A0, A1, (double*)nullptr, st);
      }
    }
  };
  // M2M from level l into l-1 only where level l-1 has M2L pairs (coarser multipoles are not used).
  // M2M and L2L use the FP64 operators for both operators: their inputs and outputs are FP64, and the
  // FP64 matrix-core tiles ran faster here than the FP32 ones (0.115 / 0.120 ms against 0.157 / 0.161 ms
  // per matvec at 209k nodes)
  translate_levels(true, w.M0, w.M1);
  rec(3);
  if (!w.stage) {
    // no staging: the FP64 M2L tiles accumulate straight into the locals (atomics, as M2M and L2L do)
    fill(c, L0, nbC, 0., st);
    fill(c, L1, nbC, 0., st);
    launch_translate(c, c->m2l, c->ops.p, w.M0, w.M1, L0, L1, (double*)nullptr, st);
    rec(4);
  } else if (lo) {
    float* stf = reinterpret_cast<float*>(w.stage);
    if (c->m2l_rot) launch_m2l_rot(c, st, stf, w.M0, w.M1, w.m32);
    else launch_translate(c, c->m2l_pw_on ? c->m2l_pw : c->m2l, c->ops32.p, w.M0, w.M1, nullptr, nullptr, stf, st);
    rec(4);
    const int Cw = fp32_coeffs(c);
    dev::m2l_reduce(st, Cw, Cw, c->n_boxes, c->m2l_rorder.p, c->m2l_tptr.p, c->m2l_tpairs.p, (const float*)stf, L0,
                    L1);
  } else {
    double* std_ = reinterpret_cast<double*>(w.stage);
    launch_translate(c, c->m2l, c->ops.p, w.M0, w.M1, nullptr, nullptr, std_, st);
    rec(4);
    dev::m2l_reduce(st, c->C, c->n_boxes, c->m2l_rorder.p, c->m2l_tptr.p, c->m2l_tpairs.p, (const double*)std_, L0,
                    L1);
  }
  rec(5);
  // L2L into level l only from parents at levels with M2L pairs (the locals of coarser levels are zero)
  translate_levels(false, L0, L1);
  rec(6);
  if (split) HIPCHECK(hipStreamWaitEvent(st, w.join, 0));
  const int part = split ? 2 : 0;
  if (lo) launch_gemv_u(c, in, L0, L1, w.y0, part, st);
  else launch_gemv(c, c->near, c->near.mat.p, w.buf0, w.buf1, w.y0, w.y1, part, st);
  rec(7);
  dev::assemble(st, mode, n, c->kind.p, c->point.p, c->alpha.p, phi, w.y0, lo ? nullptr : w.y1, c->phi_k.p, out);
  rec(8);
  HIPCHECK(hipGetLastError());
}

// (graph teardown releases the runtime's device resources of the graph: kept off other threads' captures
// like the library's own frees)
void drop_graphs(bemfmm_ctx* c) {
  std::shared_lock<std::shared_mutex> gate(g_capture_gate);
  for (auto& g : c->gexec)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

// mode 0 (A x) is replayed from a HIP graph captured once per context and precision on the context's own
// input/output buffers (about 20 kernels per matvec become one launch; no memset nodes: the zero fills
// are kernels). Event timing (bemfmm_matvec_profile) uses the eager launches. The copy into the graph's
// input is skipped when gin already equals `in` (scale_basis wrote both); out == gout skips the copy out.
bool graph_mode(const bemfmm_ctx* c) { return c->opts.use_graphs && !c->prof; }
void evaluate(bemfmm_ctx* c, int mode, const double* in, double* out, bool lo = false) {
  c->n_matvec++;
  const double* held = c->gin_holds;
  c->gin_holds = nullptr;
  if (mode != 0 || !graph_mode(c)) {
    evaluate_launch(c, mode, in, out, lo, context_work(c));
    return;
  }
  const int gi = lo ? 1 : 0;
  const size_t bytes = (size_t)c->n_dof * sizeof(double);
  if (!c->gexec[gi]) {
    // no device-wide synchronisation (frees of other contexts) while this stream is being captured
    std::unique_lock<std::shared_mutex> gate(g_capture_gate);
    hipGraph_t g = nullptr;
    MatvecWork w = context_work(c);
    w.stream = c->gstream;
    w.stream2 = c->gstream2;
    w.fork = c->ev_gfork;
    w.join = c->ev_gjoin;
    HIPCHECK(hipStreamBeginCapture(c->gstream, hipStreamCaptureModeThreadLocal));
    try {
      evaluate_launch(c, 0, c->gin.p, c->gout.p, lo, w);
    } catch (...) {
      (void)hipStreamEndCapture(c->gstream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIPCHECK(hipStreamEndCapture(c->gstream, &g));
    const hipError_t e = hipGraphInstantiate(&c->gexec[gi], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHECK(e);
  }
  if (in != c->gin.p && in != held)
    HIPCHECK(hipMemcpyAsync(c->gin.p, in, bytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCHECK(hipEventRecord(c->ev_gin, c->stream));
  HIPCHECK(hipStreamWaitEvent(c->gstream, c->ev_gin, 0));
  HIPCHECK(hipGraphLaunch(c->gexec[gi], c->gstream));
  HIPCHECK(hipEventRecord(c->ev_gout, c->gstream));
  HIPCHECK(hipStreamWaitEvent(c->stream, c->ev_gout, 0));
  if (out != c->gout.p) HIPCHECK(hipMemcpyAsync(out, c->gout.p, bytes, hipMemcpyDeviceToDevice, c->stream));
}

// Internal DOF order (solveGMRES reindexes DOFs by Morton key, GMRES_FMM.hpp:3164-3246): keys of the DOFs'
// points on the GPU, stable LSD radix sort (hipCUB) of the 63-bit keys, so equal keys (the DOFs of one
// point) keep the caller's order. opts.dof_reorder == 0 keeps the caller's numbering.
void dof_order(bemfmm_ctx* c) {
  const int n = c->n_dof;
  // indices are used by the key kernel and the renumbering below: reject bad input before either
  for (int i = 0; i < n; ++i)
    if (c->dof_point[i] < 0 || c->dof_point[i] >= c->n_nodes || c->dof_kind[i] < 0 || c->dof_kind[i] > 3)
      throw std::runtime_error("bemfmm_prepare: dof_point or dof_kind out of range");
  // quadratic faces: an edge midpoint may carry no DOF (-1), its weight then goes to the edge's two vertices
  const size_t npf = c->opts.element_order == 2 ? 6 : 3;
  for (size_t k = 0; k < c->face_dof.size(); ++k) {
    const int32_t d = c->face_dof[k];
    if (d == -1 && k % npf >= 3) continue;
    if (d < 0 || d >= n) throw std::runtime_error("bemfmm_prepare: face_dof out of range");
  }
  for (int32_t v : c->tri)
    if (v < 0 || v >= c->n_nodes) throw std::runtime_error("bemfmm_prepare: tri out of range");
  c->perm.resize(n);
  if (c->opts.dof_reorder) {
    hipStream_t s = c->stream;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int p = 0; p < c->n_nodes; ++p)
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], c->X[3 * p + a]);
        hi[a] = std::max(hi[a], c->X[3 * p + a]);
      }
    auto& dX = c->sb.dX;
    auto &dpt = c->sb.dpt, &v0 = c->sb.v0, &v1 = c->sb.v1;
    auto &k0 = c->sb.k0, &k1 = c->sb.k1;
    dX.upload(c->X, s);
    dpt.upload(c->dof_point, s);
    v0.alloc(n);
    v1.alloc(n);
    k0.alloc(n);
    k1.alloc(n);
    const double bounds[6] = {lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]};
    dev::morton_keys(s, n, dpt.p, dX.p, bounds, k0.p, v0.p);
    HIPCHECK(hipGetLastError());
    size_t tmp_bytes = 0;
    HIPCHECK(dev::radix_sort_pairs(s, nullptr, tmp_bytes, k0.p, k1.p, v0.p, v1.p, n, 63));
    auto& tmp = c->sb.sort_tmp;
    tmp.alloc(tmp_bytes);
    HIPCHECK(dev::radix_sort_pairs(s, tmp.p, tmp_bytes, k0.p, k1.p, v0.p, v1.p, n, 63));
    HIPCHECK(hipMemcpyAsync(c->perm.data(), v1.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
  } else {
    for (int i = 0; i < n; ++i) c->perm[i] = i;
  }
  c->inv.assign(n, -1);
  c->perm_identity = true;
  for (int i = 0; i < n; ++i) {
    const int j = c->perm[i];
    if (j < 0 || j >= n || c->inv[j] >= 0) throw std::runtime_error("internal: DOF order is not a permutation");
    c->inv[j] = i;
    if (j != i) c->perm_identity = false;
  }
  c->face_dof_i.resize(c->face_dof.size());
  for (size_t k = 0; k < c->face_dof.size(); ++k)
    c->face_dof_i[k] = c->face_dof[k] < 0 ? -1 : c->inv[c->face_dof[k]];  // -1: edge midpoint without a DOF
  c->kind_i.resize(n);
  c->point_i.resize(n);
  for (int i = 0; i < n; ++i) {
    c->kind_i[i] = c->dof_kind[c->perm[i]];
    c->point_i[i] = c->dof_point[c->perm[i]];
  }
  c->d_perm.upload(c->perm, c->stream);
}

// caller-ordered host vector <-> internal order
void host_to_internal(const bemfmm_ctx* c, const double* in, double* out) {
  for (int i = 0; i < c->n_dof; ++i) out[i] = in[c->perm[i]];
}
void host_to_caller(const bemfmm_ctx* c, const double* in, double* out) {
  for (int i = 0; i < c->n_dof; ++i) out[c->perm[i]] = in[i];
}

// knowns in internal order, Neumann phi_n scaled by L (GMRES_FMM.hpp:3076-3091)
void upload_knowns(bemfmm_ctx* c) {
  const int n = c->n_dof;
  std::vector<double> phik(n), phink(n);
  for (int i = 0; i < n; ++i) {
    const int j = c->perm[i];
    phik[i] = c->phi_bc[j];
    phink[i] = (c->kind_i[i] != 0) ? c->phin_bc[j] * c->plan->Lscale : c->phin_bc[j];
  }
  // ROBIN: phi_n = kappa phi + g, so in scaled coordinates (phi_n' = L phi_n) kappa' = L kappa
  std::vector<double> kap(n, 0.);
  if (!c->robin_bc.empty())
    for (int i = 0; i < n; ++i)
      if (c->kind_i[i] == 3) kap[i] = c->robin_bc[c->perm[i]] * c->plan->Lscale;
  if (c->phi_k.n != (size_t)n) c->phi_k.alloc(n);
  if (c->phin_k.n != (size_t)n) c->phin_k.alloc(n);
  if (c->robin.n != (size_t)n) c->robin.alloc(n);
  HIPCHECK(hipMemcpyAsync(c->phi_k.p, phik.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipMemcpyAsync(c->phin_k.p, phink.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipMemcpyAsync(c->robin.p, kap.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
}

// Maps the plan's operators to slots of the context's device table [slot][KP][RP] (zero padding), uploading
// only operators the table does not hold yet: a re-prepare after nodes moved (RK stages of a session) reuses
// nearly all of them. The table grows by doubling (old slots copied on the device).
// pw_kind != 0: the M2L offsets take that cache kind (plane-wave composed operators) instead of SimpleM2L's.
std::vector<int32_t> update_operator_table(bemfmm_ctx* c, const Plan& pl, int pw_kind = 0) {
  auto key_of = [&](int op) {
    auto k = bemfmm::op_key(pl, op);
    if (pw_kind && k[0] == 2) k[0] = pw_kind;
    return k;
  };
  hipStream_t s = c->stream;
  const int C = pl.C, KP = c->KP, RP = c->RP;
  const int64_t per = (int64_t)KP * RP;
  if (c->ops_P != pl.P) {
    c->op_slot.clear();
    c->ops_used = 0;
    c->ops_cap = 0;
    c->ops.release();
    c->ops32.release();
    c->ops32_valid = 0;
    c->ops_P = pl.P;
  }
  std::vector<int32_t> slot(pl.n_ops);
  std::vector<int> fresh;  // plan ops to upload
  for (int op = 0; op < pl.n_ops; ++op) {
    const auto key = key_of(op);
    auto it = c->op_slot.find(key);
    if (it == c->op_slot.end()) {
      it = c->op_slot.emplace(key, (int32_t)(c->ops_used + fresh.size())).first;
      fresh.push_back(op);
    }
    slot[op] = it->second;
  }
  const int need = c->ops_used + (int)fresh.size();
  if (need > c->ops_cap) {
    const int cap = std::max(need, std::max(64, 2 * c->ops_cap));
    DBuf<double> t;
    t.alloc((size_t)cap * per);
    if (c->ops_used > 0)
      HIPCHECK(hipMemcpyAsync(t.p, c->ops.p, (size_t)c->ops_used * per * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIPCHECK(hipStreamSynchronize(s));
    c->ops = std::move(t);
    c->ops_cap = cap;
  }
  // operators not in the process-wide cache yet are computed in parallel first
  {
    std::vector<std::array<int, 4>> keys;
    keys.reserve(fresh.size());
    for (int op : fresh) keys.push_back(key_of(op));
    bemfmm::warm_operator_cache(pl.P, keys);
  }
  // upload the new operators in batches through pinned staging, padded to [KP][RP]
  constexpr int kBatch = 32;
  if (!fresh.empty() && c->op_stage_n < (size_t)kBatch * per) {
    if (c->op_stage) (void)pinned_free(c->op_stage);
    c->op_stage = nullptr;
    HIPCHECK(pinned_malloc((void**)&c->op_stage, (size_t)kBatch * per * sizeof(double)));
    c->op_stage_n = (size_t)kBatch * per;
  }
  for (size_t b0 = 0; b0 < fresh.size(); b0 += kBatch) {
    const int nb = (int)std::min<size_t>(kBatch, fresh.size() - b0);
    for (int i = 0; i < nb; ++i) {
      const auto key = key_of(fresh[b0 + i]);
      const auto v = bemfmm::cached_operator(pl.P, key[0], key[1], key[2], key[3]);
      double* dst = c->op_stage + (size_t)i * per;
      std::fill(dst, dst + per, 0.);
      for (int k = 0; k < C; ++k) std::memcpy(dst + (size_t)k * RP, v->data() + (size_t)k * C, C * sizeof(double));
    }
    HIPCHECK(hipMemcpyAsync(c->ops.p + (size_t)(c->ops_used + b0) * per, c->op_stage, (size_t)nb * per * sizeof(double),
                            hipMemcpyHostToDevice, s));
    HIPCHECK(hipStreamSynchronize(s));  // the staging is refilled by the next batch
  }
  c->ops_used = need;
  c->ops_uploaded = (int)fresh.size();
  return slot;
}

// Rotation factors of order P of the plan's M2L offsets in the FP32 table R [slot][RS], indexed by the
// operator table's slots; only slots without factors yet are computed (process-wide cache) and uploaded.
void update_rotation_table(bemfmm_ctx* c, const Plan& pl, const std::vector<int32_t>& slot, bemfmm_ctx::RotTable& R,
                           int P) {
  hipStream_t s = c->stream;
  const int RS = dev::rot_size(P);  // padded device layout
  if (R.RS != RS) {
    R.f.release();
    R.have.clear();
    R.RS = RS;
  }
  if ((int64_t)R.have.size() < c->ops_cap) {
    DBuf<float> t;
    t.alloc((size_t)c->ops_cap * RS);
    if (!R.have.empty())
      HIPCHECK(hipMemcpyAsync(t.p, R.f.p, R.have.size() * RS * sizeof(float), hipMemcpyDeviceToDevice, s));
    HIPCHECK(hipStreamSynchronize(s));
    R.f = std::move(t);
    R.have.resize(c->ops_cap, 0);
  }
  std::vector<int> fresh;
  for (int op = 16; op < pl.n_ops; ++op)
    if (!R.have[slot[op]]) {
      R.have[slot[op]] = 1;
      fresh.push_back(op);
    }
  if (fresh.empty()) return;
  {
    std::vector<std::array<int, 4>> keys;
    for (int op : fresh) {
      const auto k = bemfmm::op_key(pl, op);
      keys.push_back({3, k[1], k[2], k[3]});
    }
    bemfmm::warm_operator_cache(P, keys);  // factors of new offsets, computed in parallel once per process
  }
  const size_t cap = c->op_stage_n * 2 / RS;  // rotations per pinned staging batch (the staging holds doubles)
  if (cap < 1) {
    if (c->op_stage) (void)pinned_free(c->op_stage);
    c->op_stage = nullptr;
    c->op_stage_n = (size_t)RS * 64;
    HIPCHECK(pinned_malloc((void**)&c->op_stage, c->op_stage_n * sizeof(double)));
  }
  const size_t batch = c->op_stage_n * 2 / RS;
  float* stg = reinterpret_cast<float*>(c->op_stage);
  for (size_t b0 = 0; b0 < fresh.size(); b0 += batch) {
    const size_t nb = std::min(batch, fresh.size() - b0);
    // consecutive slots are uploaded together: runs of fresh ops with consecutive slots
    for (size_t i = 0; i < nb; ++i) {
      const auto key = bemfmm::op_key(pl, fresh[b0 + i]);
      const auto v = bemfmm::cached_operator(P, 3, key[1], key[2], key[3]);
      dev::rot_pad(P, v->data(), stg + i * RS);
    }
    size_t i = 0;
    while (i < nb) {
      size_t j = i + 1;
      while (j < nb && slot[fresh[b0 + j]] == slot[fresh[b0 + j - 1]] + 1) ++j;
      HIPCHECK(hipMemcpyAsync(R.f.p + (size_t)slot[fresh[b0 + i]] * RS, stg + i * RS, (j - i) * RS * sizeof(float),
                              hipMemcpyHostToDevice, s));
      i = j;
    }
    HIPCHECK(hipStreamSynchronize(s));  // the staging is refilled by the next batch
  }
}

// Tile-size autotuner (opts.m2l_tile_pairs == 0): times the M2L translation kernel of this plan with
// 16 and 32 pairs per tile on the multipoles the rigid-mode evaluation just left in M (one warm-up and
// two timed launches each, HIP events on the context's stream) and keeps the faster tiling. The choice
// is kept for the process per (expansion order, size class = log2 of the pair count), so later plans of
// that class reuse it instead of timing again (and instead of timing behind other contexts' work).
// Plans with few pairs keep the default.
void tune_m2l_tiles(bemfmm_ctx* c, const Plan& pl) {
  c->m2l_tp = c->m2l.tp;
  if (c->opts.m2l_tile_pairs != 0 || pl.n_m2l_pairs < 20000) return;
  static std::mutex mu;
  static std::map<std::pair<int, int>, int> chosen;
  const std::pair<int, int> cls{pl.P, (int)std::log2((double)pl.n_m2l_pairs)};
  int known = 0;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = chosen.find(cls);
    if (it != chosen.end()) known = it->second;
  }
  if (known == c->m2l.tp) return;
  if (known) {
    DTiles& alt = c->sb.alt_tiles;
    alt.upload(pl.m2l, c->stream, c->m2l_slot, known);
    std::swap(c->m2l, alt);  // the previous tiling stays allocated for the next re-prepare
    c->m2l_tp = known;
    return;
  }
  hipStream_t s = c->stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  struct Events {  // destroyed on every exit, including a throw from a launch below
    hipEvent_t* e[2];
    ~Events() {
      for (auto* x : e)
        if (*x) (void)hipEventDestroy(*x);
    }
  } guard{{&e0, &e1}};
  HIPCHECK(hipEventCreate(&e0));
  HIPCHECK(hipEventCreate(&e1));
  auto time_tiles = [&](DTiles& t) {
    float best = 1e30f;
    for (int r = 0; r < 3; ++r) {
      HIPCHECK(hipEventRecord(e0, s));
      launch_translate(c, t, c->ops.p, c->M0.p, c->M1.p, nullptr, nullptr, c->m2l_stage.p);
      HIPCHECK(hipEventRecord(e1, s));
      HIPCHECK(hipEventSynchronize(e1));
      float ms = 0.f;
      HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
      if (r > 0) best = std::min(best, ms);
    }
    return best;
  };
  DTiles& alt = c->sb.alt_tiles;
  alt.upload(pl.m2l, s, c->m2l_slot, c->m2l.tp == 16 ? 32 : 16);
  const float t_cur = time_tiles(c->m2l), t_alt = time_tiles(alt);
  if (t_alt < t_cur) std::swap(c->m2l, alt);
  c->m2l_tp = c->m2l.tp;
  {
    std::lock_guard<std::mutex> lk(mu);
    chosen.emplace(cls, c->m2l_tp);
  }
  HIPCHECK(hipGetLastError());
}

// Launch shapes of the near-field GEMV (rows per wave: 4, 2, 8 for variants 0, 1, 2) and of P2M (loads in flight:
// 8, 4, 16), per operator precision: as given in the options, or timed with HIP events (best of two runs after a
// warm-up) at the first prepare of a problem class (precision, expansion order, near-field entries and P2M matrix
// size to within a factor of two) and reused by later contexts of that class. The timed runs read the context's
// work vectors, cleared first, and write work buffers that every matvec rewrites.
constexpr int kRowsPerWave[3] = {4, 2, 8}, kP2MLoads[3] = {8, 4, 16};

void tune_launch_shapes(bemfmm_ctx* c, const Plan& pl) {
  auto variant_of = [](const int* tab, int v) {
    for (int k = 0; k < 3; ++k)
      if (tab[k] == v) return k;
    return 0;
  };
  const bool fp32 = c->opts.precision == BEMFMM_PRECISION_FP32_REFINE;
  for (int k = 0; k < 2; ++k) {
    c->gemv_var[k] = variant_of(kRowsPerWave, c->opts.near_rows_per_wave);
    c->p2m_var[k] = variant_of(kP2MLoads, c->opts.p2m_loads);
  }
  const bool tune_gemv = c->opts.near_rows_per_wave == 0 && pl.near_nnz >= (int64_t(1) << 21) && c->near.nblk > 0;
  const bool tune_p2m = c->opts.p2m_loads == 0 && c->p2m.nblk >= 256;
  if (!tune_gemv && !tune_p2m) return;
  static std::mutex mu;
  static std::map<std::array<int, 4>, std::array<int, 4>> chosen;  // class -> {gemv FP64, FP32, p2m FP64, FP32}
  const std::array<int, 4> cls{fp32 ? 1 : 0, pl.P, (int)std::log2((double)std::max<int64_t>(pl.near_nnz, 1)),
                               (int)std::log2((double)std::max<size_t>(c->p2m.mat.n, 1))};
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = chosen.find(cls);
    if (it != chosen.end()) {
      if (tune_gemv) c->gemv_var[0] = it->second[0], c->gemv_var[1] = it->second[1];
      if (tune_p2m) c->p2m_var[0] = it->second[2], c->p2m_var[1] = it->second[3];
      return;
    }
  }
  hipStream_t s = c->stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  struct Events {  // destroyed on every exit, including a throw from a launch below
    hipEvent_t* e[2];
    ~Events() {
      for (auto* x : e)
        if (*x) (void)hipEventDestroy(*x);
    }
  } guard{{&e0, &e1}};
  HIPCHECK(hipEventCreate(&e0));
  HIPCHECK(hipEventCreate(&e1));
  const int64_t nvec = (int64_t)c->n_dof + (int64_t)pl.n_boxes * pl.C;
  fill(c, c->buf0.p, nvec, 0.);
  fill(c, c->buf1.p, nvec, 0.);
  auto best_of = [&](auto&& launch) {
    float best = 1e30f;
    for (int r = 0; r < 3; ++r) {
      HIPCHECK(hipEventRecord(e0, s));
      launch();
      HIPCHECK(hipEventRecord(e1, s));
      HIPCHECK(hipEventSynchronize(e1));
      float ms = 0.f;
      HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
      if (r > 0) best = std::min(best, ms);
    }
    return best;
  };
  for (int prec = 0; prec < (fp32 ? 2 : 1); ++prec) {
    if (tune_gemv) {
      float bt = 1e30f;
      int bv = 0;
      for (int v = 0; v < 3; ++v) {
        c->gemv_var[prec] = v;
        const float t = best_of([&] {
          if (prec) launch_gemv_u(c, c->buf0.p, c->buf0.p + c->n_dof, c->buf1.p + c->n_dof, c->y0.p, 0, s);
          else launch_gemv(c, c->near, c->near.mat.p, c->buf0.p, c->buf1.p, c->y0.p, c->y1.p);
        });
        if (t < bt) bt = t, bv = v;
      }
      c->gemv_var[prec] = bv;
    }
    if (tune_p2m) {
      float bt = 1e30f;
      int bv = 0;
      for (int v = 0; v < 3; ++v) {
        c->p2m_var[prec] = v;
        const float t = best_of([&] {
          if (prec) launch_p2m(c, c->p2m_mat32.p, c->buf0.p, c->buf1.p, c->M0.p, c->M1.p, s);
          else launch_p2m(c, c->p2m.mat.p, c->buf0.p, c->buf1.p, c->M0.p, c->M1.p, s);
        });
        if (t < bt) bt = t, bv = v;
      }
      c->p2m_var[prec] = bv;
    }
  }
  HIPCHECK(hipGetLastError());
  std::lock_guard<std::mutex> lk(mu);
  chosen.emplace(cls, std::array<int, 4>{c->gemv_var[0], c->gemv_var[1], c->p2m_var[0], c->p2m_var[1]});
}

int bits_for(uint64_t v) {
  int b = 1;
  while (b < 64 && (v >> b)) ++b;
  return b;
}

// rocPRIM calls of the setup on the context's stream, with their temporary storage in sb.sort_tmp
void dsort_keys(bemfmm_ctx* c, const uint64_t* in, uint64_t* out, int64_t n, int end_bit) {
  if (n <= 0) return;
  size_t bytes = 0;
  HIPCHECK(dev::sort_keys_u64(c->stream, nullptr, bytes, in, out, n, 0, end_bit));
  c->sb.sort_tmp.alloc(bytes);
  HIPCHECK(dev::sort_keys_u64(c->stream, c->sb.sort_tmp.p, bytes, in, out, n, 0, end_bit));
}
void dscan_i32(bemfmm_ctx* c, bool inclusive, const int32_t* in, int32_t* out, int64_t n) {
  if (n <= 0) return;
  size_t bytes = 0;
  if (inclusive) HIPCHECK(dev::scan_inclusive_i32(c->stream, nullptr, bytes, in, out, n));
  else HIPCHECK(dev::scan_exclusive_i32(c->stream, nullptr, bytes, in, out, n));
  c->sb.sort_tmp.alloc(bytes);
  if (inclusive) HIPCHECK(dev::scan_inclusive_i32(c->stream, c->sb.sort_tmp.p, bytes, in, out, n));
  else HIPCHECK(dev::scan_exclusive_i32(c->stream, c->sb.sort_tmp.p, bytes, in, out, n));
}
void dscan_i64(bemfmm_ctx* c, const int64_t* in, int64_t* out, int64_t n) {
  if (n <= 0) return;
  size_t bytes = 0;
  HIPCHECK(dev::scan_exclusive_i64(c->stream, nullptr, bytes, in, out, n));
  c->sb.sort_tmp.alloc(bytes);
  HIPCHECK(dev::scan_exclusive_i64(c->stream, c->sb.sort_tmp.p, bytes, in, out, n));
}
template <typename T>
T dread(bemfmm_ctx* c, const T* p) {
  T v{};
  HIPCHECK(hipMemcpyAsync(&v, p, sizeof(T), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return v;
}

// The plan's tree stage on the device (tree.hip): level by level, the faces of every box that splits are keyed
// by (box, cell, face) and radix-sorted, each distinct (box, cell) becomes a child box (in depth-first order),
// and each face moves into its child if the child's inside test holds; then the faces sorted by path key give
// the depth-first face order and the box ranges, the internal offsets go up the levels, every node descends
// to its deepest box, and the near / M2L lists are built level-synchronously from the root's children pairs.
// The result equals plan_tree_lists (same boxes, face order, lists and target boxes): the box arithmetic is
// shared (treemath.hpp) and compiled without FMA contraction on both sides.
void plan_tree_lists_device(bemfmm_ctx* c, Plan& pl) {
  const bemfmm::Options& opt = pl.opt;
  if (opt.max_level > 11) {  // 5-bit digits per level in 64-bit path keys
    bemfmm::plan_tree_lists(pl);
    return;
  }
  hipStream_t s = c->stream;
  auto& T = c->tb;
  const int F = pl.n_faces, N = pl.n_nodes;
  const double t0 = now();
  auto sort_keys = [&](const uint64_t* in, uint64_t* out, int64_t n, int end_bit) {
    size_t bytes = 0;
    HIPCHECK(dev::sort_keys_u64(s, nullptr, bytes, in, out, n, 0, end_bit));
    c->sb.sort_tmp.alloc(bytes);
    HIPCHECK(dev::sort_keys_u64(s, c->sb.sort_tmp.p, bytes, in, out, n, 0, end_bit));
  };
  auto scan = [&](bool inclusive, const int32_t* in, int32_t* out, int64_t n) {
    size_t bytes = 0;
    if (inclusive) HIPCHECK(dev::scan_inclusive_i32(s, nullptr, bytes, in, out, n));
    else HIPCHECK(dev::scan_exclusive_i32(s, nullptr, bytes, in, out, n));
    c->sb.sort_tmp.alloc(bytes);
    if (inclusive) HIPCHECK(dev::scan_inclusive_i32(s, c->sb.sort_tmp.p, bytes, in, out, n));
    else HIPCHECK(dev::scan_exclusive_i32(s, c->sb.sort_tmp.p, bytes, in, out, n));
  };
  auto read_err = [&]() {
    int32_t e = 0;
    HIPCHECK(hipMemcpyAsync(&e, T.err.p, sizeof(e), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    if (e & 1) throw std::runtime_error("octree box with more than 27 cells");
    if (e & 2) throw std::runtime_error("empty octree box (face on a box boundary)");
  };
  // box arrays, grown keeping the first `keep` boxes
  auto boxes = [&](size_t need, size_t keep) {
    T.geom.reserve_keep(need, keep, s);
    T.parent.reserve_keep(need, keep, s);
    T.level.reserve_keep(need, keep, s);
    T.pk.reserve_keep(need, keep, s);
    T.count.reserve_keep(need, keep, s);
    T.nchild.reserve_keep(need, keep, s);
    T.first_child.reserve_keep(need, keep, s);
    T.child_tab.reserve_keep(need * dev::kTreeCells, keep * dev::kTreeCells, s);
  };
  T.pole.upload(pl.pole, s);
  T.maxoff.upload(pl.maxoff, s);
  T.X.upload(pl.Xs, s);
  T.fbox.alloc(F);
  T.fkey.alloc(F);
  T.keys0.alloc(F);
  T.keys1.alloc(F);
  T.flags.alloc(F);
  T.ci.alloc(F);
  T.err.alloc(1);
  T.ctr.alloc(1);
  T.tot.alloc(3);
  dev::fill_i32(s, F, T.fbox.p, 0);
  dev::fill_u64(s, F, T.fkey.p, 0);
  dev::fill_i32(s, 1, T.err.p, 0);
  boxes(std::max<size_t>(64, (size_t)F / 16), 0);
  {
    const int32_t root_i[5] = {-1, 0, F, 0, -1};  // parent, level, count, nchild, first_child
    const uint64_t zero = 0;
    std::vector<int32_t> tab(dev::kTreeCells, -1);
    HIPCHECK(hipMemcpyAsync(T.geom.p, &pl.root, sizeof(BoxGeom), hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(T.parent.p, &root_i[0], 4, hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(T.level.p, &root_i[1], 4, hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(T.count.p, &root_i[2], 4, hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(T.nchild.p, &root_i[3], 4, hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(T.first_child.p, &root_i[4], 4, hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(T.pk.p, &zero, 8, hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(T.child_tab.p, tab.data(), tab.size() * 4, hipMemcpyHostToDevice, s));
    HIPCHECK(hipStreamSynchronize(s));  // the sources are on this stack frame
  }
  std::vector<int32_t> level_ptr{0, 1};
  int digits = 0;
  for (int l = 0;; ++l) {
    const int lb = level_ptr[l], le = level_ptr[l + 1];
    dev::fill_i32(s, 1, (int32_t*)T.ctr.p, 0);
    dev::tree_split(s, F, lb, opt.leaf_max, l < opt.max_level ? 1 : 0, T.pole.p, T.fbox.p, T.fkey.p, T.geom.p,
                    T.count.p, T.keys0.p, T.ctr.p, T.err.p);
    HIPCHECK(hipGetLastError());
    ++digits;
    uint32_t ne = 0;
    HIPCHECK(hipMemcpyAsync(&ne, T.ctr.p, sizeof(ne), hipMemcpyDeviceToHost, s));
    read_err();
    if (ne == 0) break;
    sort_keys(T.keys0.p, T.keys1.p, ne, 32 + bits_for((uint64_t)(le - lb) * 32));
    dev::tree_heads(s, (int)ne, T.keys1.p, T.flags.p);
    scan(true, T.flags.p, T.ci.p, ne);
    int32_t nc = 0;
    HIPCHECK(hipMemcpyAsync(&nc, T.ci.p + (ne - 1), sizeof(nc), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    boxes((size_t)le + nc, (size_t)le);
    dev::tree_children(s, (int)ne, T.keys1.p, T.ci.p, lb, le, T.geom.p, T.parent.p, T.level.p, T.pk.p, T.count.p,
                       T.nchild.p, T.first_child.p, T.child_tab.p);
    dev::tree_assign(s, (int)ne, T.keys1.p, T.ci.p, le, T.pole.p, T.geom.p, T.fbox.p, T.fkey.p, T.count.p);
    dev::tree_nonempty(s, nc, le, T.count.p, T.err.p);
    HIPCHECK(hipGetLastError());
    read_err();
    level_ptr.push_back(le + nc);
  }
  const int nl = (int)level_ptr.size() - 1, nb = level_ptr.back();
  // depth-first face order and box ranges
  T.vals.alloc(F);
  T.perm.alloc(F);
  T.range.alloc(2 * (size_t)nb);
  T.off.alloc(nb);
  T.target.alloc(N);
  dev::iota_i32(s, F, T.vals.p);
  {
    size_t bytes = 0;
    HIPCHECK(dev::sort_pairs_u64_i32(s, nullptr, bytes, T.fkey.p, T.keys1.p, T.vals.p, T.perm.p, F, 0, 5 * digits));
    c->sb.sort_tmp.alloc(bytes);
    HIPCHECK(dev::sort_pairs_u64_i32(s, c->sb.sort_tmp.p, bytes, T.fkey.p, T.keys1.p, T.vals.p, T.perm.p, F, 0,
                                     5 * digits));
  }
  dev::tree_ranges(s, nb, digits, T.level.p, T.pk.p, T.keys1.p, F, T.range.p);
  dev::tree_leaf_off(s, nb, T.nchild.p, T.range.p, T.perm.p, T.maxoff.p, T.off.p);
  for (int l = nl - 2; l >= 0; --l)
    dev::tree_parent_off(s, level_ptr[l], level_ptr[l + 1], T.nchild.p, T.first_child.p, T.off.p);
  dev::tree_targets(s, N, T.X.p, T.geom.p, T.nchild.p, T.child_tab.p, T.target.p);
  HIPCHECK(hipGetLastError());
  std::vector<BoxGeom> geom;
  std::vector<int32_t> nchild, first_child;
  T.geom.download(geom, nb, s);
  T.nchild.download(nchild, nb, s);
  T.first_child.download(first_child, nb, s);
  T.parent.download(pl.box_parent, nb, s);
  T.perm.download(pl.face_perm, F, s);
  T.range.download(pl.box_range, 2 * (size_t)nb, s);
  T.target.download(pl.target_box, N, s);
  pl.n_levels = nl;
  pl.n_boxes = nb;
  pl.level_ptr = level_ptr;
  pl.box_level.resize(nb);
  pl.box_leaf.resize(nb);
  pl.box_center.resize(3 * (size_t)nb);
  pl.n_leaves = 0;
  for (int l = 0; l < nl; ++l)
    for (int b = level_ptr[l]; b < level_ptr[l + 1]; ++b) pl.box_level[b] = l;
  for (int b = 0; b < nb; ++b) {
    pl.box_leaf[b] = nchild[b] == 0 ? 1 : 0;
    pl.n_leaves += pl.box_leaf[b];
    for (int a = 0; a < 3; ++a) pl.box_center[3 * (size_t)b + a] = geom[b].X[a];
  }
  const double t1 = now();
  pl.t_tree += t1 - t0;

  // near / M2L lists, level-synchronous from the pairs of the root's children
  {
    std::vector<uint64_t> p0;
    for (int a = first_child[0]; a < first_child[0] + nchild[0]; ++a)
      for (int b = first_child[0]; b < first_child[0] + nchild[0]; ++b)
        p0.push_back(((uint64_t)(uint32_t)a << 32) | (uint32_t)b);
    T.pairs.upload(p0, s);
  }
  int64_t np = nchild[0] * (int64_t)nchild[0], n_near = 0, n_m2l = 0;
  while (np > 0) {
    if (np > INT32_MAX) throw std::runtime_error("too many near candidate pairs in one level");
    for (auto* b : {&T.nearf, &T.m2lc, &T.nextc, &T.near_at, &T.m2l_at, &T.next_at}) b->alloc(np);
    dev::list_count(s, (int)np, T.pairs.p, T.geom.p, T.off.p, T.nchild.p, T.first_child.p, opt.theta, T.nearf.p,
                    T.m2lc.p, T.nextc.p);
    scan(false, T.nearf.p, T.near_at.p, np);
    scan(false, T.m2lc.p, T.m2l_at.p, np);
    scan(false, T.nextc.p, T.next_at.p, np);
    dev::list_totals(s, (int)np, T.nearf.p, T.m2lc.p, T.nextc.p, T.near_at.p, T.m2l_at.p, T.next_at.p, T.tot.p);
    int64_t tot[3];
    HIPCHECK(hipMemcpyAsync(tot, T.tot.p, sizeof(tot), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    if (tot[1] > INT32_MAX || tot[2] > INT32_MAX) throw std::runtime_error("too many box pairs in one level");
    T.near_all.reserve_keep((size_t)(n_near + tot[0]), (size_t)n_near, s);
    T.m2l_all.reserve_keep((size_t)(n_m2l + tot[1]), (size_t)n_m2l, s);
    T.next.alloc(std::max<int64_t>(tot[2], 1));
    dev::list_write(s, (int)np, T.pairs.p, T.geom.p, T.off.p, T.nchild.p, T.first_child.p, opt.theta, T.nearf.p,
                    T.near_at.p, T.m2l_at.p, T.next_at.p, T.near_all.p + n_near, T.m2l_all.p + n_m2l, T.next.p);
    HIPCHECK(hipGetLastError());
    n_near += tot[0];
    n_m2l += tot[1];
    T.pairs.swap(T.next);
    np = tot[2];
  }
  auto to_csr = [&](DBuf<uint64_t>& all, int64_t n, DBuf<int64_t>& ptr, DBuf<int32_t>& idx, std::vector<int64_t>& hp,
                    std::vector<int32_t>& hi) {
    T.keys0.alloc(std::max<int64_t>(n, 1));
    if (n) sort_keys(all.p, T.keys0.p, n, 32 + bits_for((uint64_t)nb));
    ptr.alloc(nb + 1);
    idx.alloc(std::max<int64_t>(n, 1));
    dev::csr_from_keys(s, nb, T.keys0.p, n, ptr.p, idx.p);
    HIPCHECK(hipGetLastError());
    ptr.download(hp, nb + 1, s);
    idx.download(hi, n, s);
  };
  to_csr(T.near_all, n_near, T.near_ptr, T.near_idx, pl.near_ptr, pl.near_idx);
  to_csr(T.m2l_all, n_m2l, T.m2l_ptr, T.m2l_idx, pl.m2l_ptr, pl.m2l_idx);
  pl.t_lists = now() - t1;
}

// The near-field blocks on the device, from the tree stage's device arrays: the layout plan_finish builds on the
// host (one block per target box, split into chunks of <= 32 rows in descending cost; columns: the box's sorted
// DOFs, its C L2P columns, zero-weight padding; the box's merged face ranges; the L2P coefficients per row),
// written straight into the context's near-field buffers. Face DOFs are read in face_perm order (sb.fdof).
void near_struct_device(bemfmm_ctx* c, Plan& pl) {
  hipStream_t s = c->stream;
  auto& T = c->tb;
  auto& W = c->nbuf;
  const int N = pl.n_nodes, nb = pl.n_boxes, C = pl.C, P = pl.P;
  const double t0 = now();
  // target nodes by box
  W.k0.alloc(N);
  W.k1.alloc(N);
  dev::node_keys(s, N, T.target.p, W.k0.p);
  dsort_keys(c, W.k0.p, W.k1.p, N, 32 + bits_for((uint64_t)nb));
  W.tptr.alloc(nb + 1);
  W.tnodes.alloc(N);
  dev::csr_from_keys(s, nb, W.k1.p, N, W.tptr.p, W.tnodes.p);
  // boxes with targets
  W.flags.alloc(nb);
  W.at.alloc(nb);
  dev::ns_tbox_flags(s, nb, W.tptr.p, W.flags.p);
  dscan_i32(c, true, W.flags.p, W.at.p, nb);
  const int ntb = dread(c, W.at.p + (nb - 1));
  W.tbox.alloc(ntb);
  dev::ns_tbox_compact(s, nb, W.flags.p, W.at.p, W.tbox.p);
  // merged face ranges per target box
  W.rcnt.alloc(ntb + 1);
  W.rat.alloc(ntb + 1);
  dev::ns_range_count(s, ntb, W.tbox.p, T.parent.p, T.near_ptr.p, T.near_idx.p, T.nchild.p, W.rcnt.p);
  dscan_i32(c, false, W.rcnt.p, W.rat.p, ntb + 1);
  const int R = dread(c, W.rat.p + ntb);
  for (auto* b : {&W.rkeys, &W.rkeys2}) b->alloc(R);
  for (auto* b : {&W.rends, &W.rends2}) b->alloc(R);
  dev::ns_range_write(s, ntb, W.tbox.p, T.parent.p, T.near_ptr.p, T.near_idx.p, T.nchild.p, T.range.p, W.rat.p,
                      W.rkeys.p, W.rends.p);
  {
    size_t bytes = 0;
    const int eb = 32 + bits_for((uint64_t)ntb);
    HIPCHECK(dev::sort_pairs_u64_i32(s, nullptr, bytes, W.rkeys.p, W.rkeys2.p, W.rends.p, W.rends2.p, R, 0, eb));
    c->sb.sort_tmp.alloc(bytes);
    HIPCHECK(dev::sort_pairs_u64_i32(s, c->sb.sort_tmp.p, bytes, W.rkeys.p, W.rkeys2.p, W.rends.p, W.rends2.p, R, 0,
                                     eb));
  }
  W.mstart.alloc(R);
  W.mend.alloc(R);
  W.mcnt.alloc(ntb + 1);
  W.nfaces.alloc(ntb + 1);
  dev::fill_i32(s, 1, T.err.p, 0);
  dev::ns_range_merge(s, ntb, W.rat.p, W.rcnt.p, W.rkeys2.p, W.rends2.p, W.mstart.p, W.mend.p, W.mcnt.p, W.nfaces.p,
                      T.err.p);
  if (dread(c, T.err.p) & 4) throw std::runtime_error("overlapping near ranges");
  W.fstart.alloc(ntb + 1);
  dscan_i64(c, W.nfaces.p, W.fstart.p, ntb + 1);
  const int64_t nfc = dread(c, W.fstart.p + ntb);
  // DOF columns per target box: sorted, unique
  const int npf = pl.npf;
  const int64_t nkeys = npf * nfc;
  W.dk0.alloc(nkeys);
  W.dk1.alloc(nkeys);
  W.uflag.alloc(nkeys);
  W.uat.alloc(nkeys);
  if (nkeys > INT32_MAX) throw std::runtime_error("near field too large for the device structure build");
  dev::ns_dof_keys(s, npf, nfc, ntb, W.fstart.p, W.rat.p, W.mcnt.p, W.mstart.p, W.mend.p, c->sb.fdof.p, W.dk0.p);
  dsort_keys(c, W.dk0.p, W.dk1.p, nkeys, 32 + bits_for((uint64_t)ntb));
  dev::unique_flags(s, nkeys, W.dk1.p, W.uflag.p);
  dscan_i32(c, true, W.uflag.p, W.uat.p, nkeys);
  const int64_t nu = dread(c, W.uat.p + (nkeys - 1));
  dev::compact_u64(s, nkeys, W.dk1.p, W.uflag.p, W.uat.p, W.dk0.p);
  W.dptr.alloc(ntb + 1);
  W.dofs.alloc(nu);
  dev::csr_from_keys(s, ntb, W.dk0.p, nu, W.dptr.p, W.dofs.p);
  // chunks in descending cost
  W.ccnt.alloc(ntb + 1);
  W.cat.alloc(ntb + 1);
  dev::ns_chunk_count(s, ntb, W.tptr.p, W.tbox.p, W.ccnt.p);
  dscan_i32(c, false, W.ccnt.p, W.cat.p, ntb + 1);
  const int nch = dread(c, W.cat.p + ntb);
  if (nch >= (1 << 24)) throw std::runtime_error("too many near blocks for the device structure build");
  W.cti.alloc(nch);
  W.ck0.alloc(nch);
  W.ck1.alloc(nch);
  dev::ns_chunk_ti(s, ntb, W.cat.p, W.cti.p);
  dev::fill_i32(s, 1, T.err.p, 0);
  dev::ns_chunk_keys(s, ntb, W.tptr.p, W.tbox.p, W.cat.p, W.dptr.p, C, W.ck0.p, T.err.p);
  dsort_keys(c, W.ck0.p, W.ck1.p, nch, 46);
  if (dread(c, T.err.p) & 8) throw std::runtime_error("near block too large for the device structure build");
  // block offsets
  for (auto* b : {&W.v_rows, &W.v_cols, &W.v_mat, &W.v_rng, &W.v_nnz, &W.nnzp}) b->alloc(nch + 1);
  c->near.row_ptr.alloc(nch + 1);
  c->near.col_ptr.alloc(nch + 1);
  c->near.mat_ptr.alloc(nch + 1);
  c->sb.range_ptr.alloc(nch + 1);
  dev::ns_chunk_sizes(s, nch, W.ck1.p, W.cti.p, W.cat.p, W.tptr.p, W.tbox.p, W.dptr.p, W.mcnt.p, C, W.v_rows.p,
                      W.v_cols.p, W.v_mat.p, W.v_rng.p, W.v_nnz.p);
  dscan_i64(c, W.v_rows.p, c->near.row_ptr.p, nch + 1);
  dscan_i64(c, W.v_cols.p, c->near.col_ptr.p, nch + 1);
  dscan_i64(c, W.v_mat.p, c->near.mat_ptr.p, nch + 1);
  dscan_i64(c, W.v_rng.p, c->sb.range_ptr.p, nch + 1);
  dscan_i64(c, W.v_nnz.p, W.nnzp.p, nch + 1);
  const int64_t nrows = dread(c, c->near.row_ptr.p + nch), ncols = dread(c, c->near.col_ptr.p + nch);
  const int64_t msize = dread(c, c->near.mat_ptr.p + nch), nrng = dread(c, c->sb.range_ptr.p + nch);
  const int64_t nnz = dread(c, W.nnzp.p + nch);
  if (nrows != N) throw std::runtime_error("internal: near rows do not cover the nodes");
  c->near.row_out.alloc(nrows);
  c->sb.row_block.alloc(nrows);
  c->sb.node_row.alloc(N);
  c->near.col_idx.alloc(ncols);
  c->sb.ncol_dof.alloc(nch);
  c->sb.ranges.alloc(2 * nrng);
  c->sb.range_prefix.alloc(nrng);
  W.block_fstart.alloc(nch);
  W.face_cols.alloc(nkeys);
  dev::ns_face_cols(s, npf, nfc, ntb, W.fstart.p, W.rat.p, W.mcnt.p, W.mstart.p, W.mend.p, c->sb.fdof.p, W.dptr.p,
                    W.dofs.p, W.face_cols.p);
  dev::ns_fill(s, nch, W.ck1.p, W.cti.p, W.cat.p, W.tptr.p, W.tnodes.p, W.tbox.p, W.dptr.p, W.dofs.p, W.rat.p,
               W.mcnt.p, W.mstart.p, W.mend.p, C, pl.n_dof, c->near.row_ptr.p, c->near.col_ptr.p, c->sb.range_ptr.p,
               c->near.row_out.p, c->sb.row_block.p, c->sb.node_row.p, c->near.col_idx.p, c->sb.ncol_dof.p,
               c->sb.ranges.p, c->sb.range_prefix.p, W.fstart.p, W.block_fstart.p);
  HIPCHECK(hipGetLastError());
  c->near_ncd.alloc(nch);
  HIPCHECK(hipMemcpyAsync(c->near_ncd.p, c->sb.ncol_dof.p, nch * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
  c->near.nblk = nch;
  // L2P coefficients per row
  {
    const int NT = (P + 1) * (P + 2) / 2;
    std::vector<double> cnm(NT);
    for (int n = 0; n <= P; ++n)
      for (int m = 0; m <= n; ++m) {
        double f1 = 1, f2 = 1;
        for (int k = 2; k <= n - m; ++k) f1 *= k;
        for (int k = 2; k <= n + m; ++k) f2 *= k;
        const double v = std::sqrt(f1 * f2);
        cnm[n * (n + 1) / 2 + m] = (m & 1) ? -v : v;
      }
    W.cnm.upload(cnm, s);
    W.lside.upload(pl.level_side, s);
  }
  c->sb.l2p.alloc((size_t)nrows * C);
  dev::l2p_coef(s, nrows, P, c->near.row_out.p, T.target.p, T.X.p, T.geom.p, T.level.p, W.lside.p, W.cnm.p,
                c->sb.l2p.p);
  HIPCHECK(hipGetLastError());
  c->sb.ncol_dof.download(pl.near_ncol_dof, nch, s);
  pl.near = bemfmm::BlockGemv();
  pl.near.mat_size = msize;
  pl.near_nnz = nnz;
  c->near_rows = nrows;
  pl.t_near_struct = now() - t0;
}

void do_prepare(bemfmm_ctx* c) {
  if (!c->have_geom) throw std::runtime_error("bemfmm_prepare: geometry not set");
  if (!c->have_bc) throw std::runtime_error("bemfmm_prepare: boundary values not set");
  if (c->opts.precision != BEMFMM_PRECISION_FP64 && c->opts.precision != BEMFMM_PRECISION_FP32_REFINE)
    throw std::runtime_error("bemfmm_prepare: unknown precision mode");
  if (c->opts.m2l_tile_pairs != 0 && c->opts.m2l_tile_pairs != 16 && c->opts.m2l_tile_pairs != 32 &&
      c->opts.m2l_tile_pairs != 64)
    throw std::runtime_error("bemfmm_prepare: m2l_tile_pairs must be 0 (default), 16, 32 or 64");
  if (c->opts.inner_p != 0 &&
      (c->opts.precision != BEMFMM_PRECISION_FP32_REFINE || c->opts.inner_p >= c->opts.p ||
       !dev::m2l_rot_inner_compiled(c->opts.inner_p) || (c->opts.m2l_kernel != 0 && c->opts.m2l_kernel != 2)))
    throw std::runtime_error("bemfmm_prepare: inner_p must be 0 (the plan's order) or an order below p that the "
                             "rotation M2L runs (5, 6 or 8), with FP32_REFINE and m2l_kernel 0 or 2");
  if (c->opts.near_rows_per_wave != 0 && c->opts.near_rows_per_wave != 2 && c->opts.near_rows_per_wave != 4 &&
      c->opts.near_rows_per_wave != 8)
    throw std::runtime_error("bemfmm_prepare: near_rows_per_wave must be 0 (tuned), 2, 4 or 8");
  if (c->opts.p2m_loads != 0 && c->opts.p2m_loads != 4 && c->opts.p2m_loads != 8 && c->opts.p2m_loads != 16)
    throw std::runtime_error("bemfmm_prepare: p2m_loads must be 0 (tuned), 4, 8 or 16");
  if (c->opts.element_order != 1 && c->opts.element_order != 2)
    throw std::runtime_error("bemfmm_prepare: element_order must be 1 (linear) or 2 (true quadratic)");
  if (c->opts.element_order == 2 &&
      (c->opts.preconditioner == BEMFMM_PRECOND_ILU0 || c->opts.preconditioner == BEMFMM_PRECOND_MILU0 ||
       c->opts.preconditioner == BEMFMM_PRECOND_ILUT ||
       (c->opts.preconditioner == BEMFMM_PRECOND_SCHWARZ && c->opts.schwarz_variant == 1)))
    throw std::runtime_error("bemfmm_prepare: quadratic elements take the NONE, JACOBI or mesh-graph SCHWARZ "
                             "preconditioner (the sparse preconditioner matrix is built for linear faces)");
  if (c->opts.m2l_kernel < 0 || c->opts.m2l_kernel > 4)
    throw std::runtime_error("bemfmm_prepare: m2l_kernel must be 0 (default), 1 (matrix cores), 2 (rotation), "
                             "3 (plane waves, FP32 operator) or 4 (plane waves, both operators)");
  if (c->opts.m2l_kernel >= 3 && c->opts.plane_wave_digits != 6 && c->opts.plane_wave_digits != 10)
    throw std::runtime_error("bemfmm_prepare: plane_wave_digits must be 6 or 10");
  if (c->opts.m2l_kernel == 2 && !dev::m2l_rot_compiled(c->opts.p))
    throw std::runtime_error("bemfmm_prepare: the rotation M2L is compiled for p = 8 and 10 only");
  if (c->opts.preconditioner == BEMFMM_PRECOND_ILUT &&
      (!(c->opts.ilut_drop_tol >= 0.) || c->opts.ilut_max_entries < 0 || !(c->opts.ilut_pivot_min > 0.)))
    throw std::runtime_error("bemfmm_prepare: ILUT needs ilut_drop_tol >= 0, ilut_max_entries >= 0, ilut_pivot_min > 0");
  if (c->opts.preconditioner == BEMFMM_PRECOND_MILU0 && !(c->opts.milu_omega >= 0.))
    throw std::runtime_error("bemfmm_prepare: MILU(0) needs milu_omega >= 0");
  if ((c->opts.preconditioner == BEMFMM_PRECOND_ILU0 || c->opts.preconditioner == BEMFMM_PRECOND_MILU0 ||
       c->opts.preconditioner == BEMFMM_PRECOND_ILUT) &&
      (c->opts.ilu_neighborhood < 0 || c->opts.ilu_neighborhood > 1 || c->opts.ilu_kring < 0 || c->opts.ilu_kring > 20))
    throw std::runtime_error("bemfmm_prepare: ilu_neighborhood must be 0 (BUCKETS) or 1 (K-RING), 0 <= ilu_kring <= 20");
  if (c->opts.preconditioner == BEMFMM_PRECOND_SCHWARZ &&
      (c->opts.schwarz_max_core < 1 || c->opts.schwarz_max_block < c->opts.schwarz_max_core ||
       c->opts.schwarz_max_block > kSchwarzMaxBlock || c->opts.schwarz_overlap < 0 || c->opts.schwarz_overlap > 8))
    throw std::runtime_error("bemfmm_prepare: Schwarz block sizes need 1 <= max_core <= max_block <= 128, 0 <= overlap <= 8");
  if (c->opts.preconditioner == BEMFMM_PRECOND_SCHWARZ &&
      (c->opts.schwarz_variant < 0 || c->opts.schwarz_variant > 1 || c->opts.schwarz_core_k < 0 ||
       c->opts.schwarz_core_k > 8 || c->opts.ilu_neighborhood < 0 || c->opts.ilu_neighborhood > 1 ||
       c->opts.ilu_kring < 0 || c->opts.ilu_kring > 20))
    throw std::runtime_error("bemfmm_prepare: schwarz_variant must be 0 or 1, 0 <= schwarz_core_k <= 8");
  const double t0 = now();
  bemfmm::Options o;
  o.p = c->opts.p;
  o.theta = c->opts.theta;
  o.leaf_max = c->opts.leaf_max;
  o.max_level = c->opts.max_level;
  o.q = c->opts.q;
  o.npf = c->opts.element_order == 2 ? 6 : 3;
  o.materialize_ops = false;  // operators go from the float128 cache straight into the device table
  o.fill_p2m = false;         // the P2M matrices are built on the GPU (k_p2m_build)
  c->prepared = false;
  drop_graphs(c);  // their kernels point into the buffers reallocated below
  // the Krylov basis holds v_cap vectors of n_dof: kept when the DOF count is the same (moved nodes), else
  // dropped (re-grown by the next solve)
  if (c->basis_n != c->n_dof) {
    c->V.release();
    c->partial.release();
    c->hbuf.release();
    c->v_cap = 0;
    c->basis_n = c->n_dof;
  }
  dof_order(c);
  c->plan.reset(new Plan());
  Plan& pl = *c->plan;
  bemfmm::plan_geometry(pl, o, c->n_nodes, c->X.data(), c->n_faces, c->tri.data(), c->face_dof_i.data(), c->n_dof,
                        c->kind_i.data(), c->point_i.data());
  // the tree stage and the near-field structure on the GPU (path keys use 5 bits per level of a 64-bit key)
  const bool dev_plan = c->opts.tree_on_gpu && o.max_level <= 11;
  // With the device plan, the host's share of the plan (P2M source lists, translation pair groups) runs on a
  // host thread beside the device near structure and the near-field quadrature, which do not read it; it is
  // joined before its first use below. (The future's destructor waits for it if an exception leaves first.)
  std::future<void> host_finish;
  if (dev_plan) {
    plan_tree_lists_device(c, pl);
    bemfmm::plan_level_sides(pl);
    host_finish = std::async(std::launch::async, [&pl] { bemfmm::plan_finish(pl, false); });
  } else {
    bemfmm::plan_tree_lists(pl);
    bemfmm::plan_finish(pl, true);
  }
  c->C = pl.C;
  c->n_boxes = pl.n_boxes;
  if ((pl.C + 63) / 64 > 7) throw std::runtime_error("expansion order too large for the translation kernel");
  if (2 * pl.C > 256) throw std::runtime_error("expansion order too large for the M2L reduction");
  hipStream_t s = c->stream;
  const double t1 = now();
  // ---- upload: DOF data, work vectors, face arrays, near structure ----
  if (!dev_plan) c->near.upload(pl.near, s);
  c->kind.upload(pl.dof_kind, s);
  c->point.upload(pl.dof_point, s);
  upload_knowns(c);
  const int64_t nvec = (int64_t)c->n_dof + (int64_t)pl.n_boxes * pl.C;
  c->buf0.alloc(nvec);
  c->buf1.alloc(nvec);
  c->M0.alloc((size_t)pl.n_boxes * pl.C);
  c->M1.alloc((size_t)pl.n_boxes * pl.C);
  c->y0.alloc(c->n_nodes);
  c->y1.alloc(c->n_nodes);
  c->alpha.alloc(c->n_dof);
  for (auto* b : {&c->xin, &c->yout, &c->bvec, &c->tmpv, &c->xsol, &c->xbase, &c->jacobi, &c->rvec, &c->dvec, &c->ptmp,
                 &c->gin, &c->gout})
    b->alloc(c->n_dof);
  // quadrature inputs
  auto &row_block = c->sb.row_block, &ncol_dof = c->sb.ncol_dof, &face_perm = c->sb.face_perm, &fv = c->sb.fv,
       &fdof = c->sb.fdof;
  auto &range_ptr = c->sb.range_ptr, &ranges = c->sb.ranges, &range_prefix = c->sb.range_prefix;
  auto &Xs = c->sb.Xs, &fX = c->sb.fX, &fcent = c->sb.fcent, &fcr = c->sb.fcr, &fJ = c->sb.fJ, &l2p = c->sb.l2p,
       &d_dun = c->sb.d_dun;
  face_perm.upload(pl.face_perm, s);
  // per-face arrays on the device in the depth-first face order (position in face_perm): the faces of a box
  // and of a near range are consecutive, so the quadrature and P2M kernels read them coalesced
  const int npf = pl.npf;
  std::vector<int32_t> fv_p((size_t)npf * c->n_faces), fdof_p((size_t)npf * c->n_faces);
  std::vector<double> fX_p((size_t)3 * npf * c->n_faces), fcent_p(3 * (size_t)c->n_faces), fcr_p(3 * (size_t)c->n_faces),
      fJ_p(c->n_faces);
  std::vector<int32_t> face_pos(c->n_faces);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < c->n_faces; ++k) {
    const size_t f = (size_t)pl.face_perm[k];
    face_pos[f] = k;
    for (int a = 0; a < npf; ++a) {
      fv_p[(size_t)npf * k + a] = pl.fv[npf * f + a];
      fdof_p[(size_t)npf * k + a] = pl.fdof[npf * f + a];
    }
    for (int a = 0; a < 3; ++a) {
      fcent_p[3 * (size_t)k + a] = pl.fcent[3 * f + a];
      fcr_p[3 * (size_t)k + a] = pl.fcr[3 * f + a];
    }
    for (int a = 0; a < 3 * npf; ++a) fX_p[(size_t)3 * npf * k + a] = pl.fX[3 * npf * f + a];
    fJ_p[k] = pl.fJ[f];
  }
  fv.upload(fv_p, s);
  fdof.upload(fdof_p, s);
  if (dev_plan) {
    near_struct_device(c, pl);
  } else {
    std::vector<int32_t> rb(pl.near.row_out.size());
    for (int b = 0; b < pl.near.n_blocks(); ++b)
      for (int64_t r = pl.near.row_ptr[b]; r < pl.near.row_ptr[b + 1]; ++r) rb[r] = b;
    row_block.upload(rb, s);
    ncol_dof.upload(pl.near_ncol_dof, s);
    c->near_ncd.upload(pl.near_ncol_dof, s);
    range_ptr.upload(pl.near_range_ptr, s);
    ranges.upload(pl.near_ranges, s);
    range_prefix.upload(pl.near_range_prefix, s);
    l2p.upload(pl.l2p_coef, s);
    std::vector<int64_t> node_row(c->n_nodes, -1);
    for (size_t r = 0; r < pl.near.row_out.size(); ++r) node_row[pl.near.row_out[r]] = (int64_t)r;
    c->sb.node_row.upload(node_row, s);
    c->near_rows = (int64_t)pl.near.row_out.size();
  }
  c->near.mat.alloc(pl.near.mat_size);
  Xs.upload(pl.Xs, s);
  fX.upload(fX_p, s);
  fcent.upload(fcent_p, s);
  fcr.upload(fcr_p, s);
  fJ.upload(fJ_p, s);
  {
    double dun[7][4] = {};
    static const double d1[1][4] = {{1. / 3., 1. / 3., 1. / 3., 0.5}};
    static const double d3[3][4] = {{2. / 3., 1. / 6., 1. / 6., 1. / 6.}, {1. / 6., 2. / 3., 1. / 6., 1. / 6.}, {1. / 6., 1. / 6., 2. / 3., 1. / 6.}};
    static const double d6[6][4] = {{0.816847572980459, 0.091576213509771, 0.091576213509771, 0.054975871827661},
                                    {0.091576213509771, 0.816847572980459, 0.091576213509771, 0.054975871827661},
                                    {0.091576213509771, 0.091576213509771, 0.816847572980459, 0.054975871827661},
                                    {0.108103018168070, 0.445948490915965, 0.445948490915965, 0.111690794839006},
                                    {0.445948490915965, 0.108103018168070, 0.445948490915965, 0.111690794839006},
                                    {0.445948490915965, 0.445948490915965, 0.108103018168070, 0.111690794839006}};
    static const double d7[7][4] = {{1. / 3., 1. / 3., 1. / 3., 0.1125},
                                    {0.059715871789770, 0.470142064105115, 0.470142064105115, 0.066197076394253},
                                    {0.470142064105115, 0.059715871789770, 0.470142064105115, 0.066197076394253},
                                    {0.470142064105115, 0.470142064105115, 0.059715871789770, 0.066197076394253},
                                    {0.797426985353087, 0.101286507323456, 0.101286507323456, 0.062969590272414},
                                    {0.101286507323456, 0.797426985353087, 0.101286507323456, 0.062969590272414},
                                    {0.101286507323456, 0.101286507323456, 0.797426985353087, 0.062969590272414}};
    const double(*src)[4] = o.q == 1 ? d1 : o.q == 3 ? d3 : o.q == 6 ? d6 : d7;
    for (int g = 0; g < o.q; ++g)
      for (int a = 0; a < 4; ++a) dun[g][a] = src[g][a];
    d_dun.upload(std::vector<double>(&dun[0][0], &dun[0][0] + 28), s);
    c->sb.d_dun6.upload(std::vector<double>(&d6[0][0], &d6[0][0] + 24), s);  // far rule of quadratic faces
  }
  // ---- near-field quadrature on the GPU (timed by events: the host goes on while it runs) ----
  HIPCHECK(hipEventRecord(c->prep_ev[0], s));
  fill(c, c->near.mat.p, pl.near.mat_size, 0.);
  const int64_t nrows = c->near_rows;
  if (nrows > 0) {
    dev::NearQuad a;
    a.C = pl.C;
    a.q = o.q;
    a.npf = pl.npf;
    a.dun6_rule = c->sb.d_dun6.p;
    for (int32_t v : pl.near_ncol_dof) a.max_ncd = std::max(a.max_ncd, (int)v);
    if (dev_plan) {
      a.face_cols = c->nbuf.face_cols.p;
      a.block_fstart = c->nbuf.block_fstart.p;
    }
    a.near_region = pl.near_region;
    a.row_block = row_block.p;
    a.ncol_dof = ncol_dof.p;
    a.range_ptr = range_ptr.p;
    a.ranges = ranges.p;
    a.range_prefix = range_prefix.p;
    a.face_perm = face_perm.p;
    a.Xs = Xs.p;
    a.fX = fX.p;
    a.fcent = fcent.p;
    a.fcr = fcr.p;
    a.fJ = fJ.p;
    a.fv = fv.p;
    a.fdof = fdof.p;
    a.l2p = l2p.p;
    a.dun_rule = d_dun.p;
    dev::near_quad(s, nrows, a, c->near.view(), c->near.mat.p);
  }
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipEventRecord(c->prep_ev[1], s));
  // ---- the host's share of the plan, then the P2M blocks and the translation operators ----
  if (host_finish.valid()) host_finish.get();
  c->p2m_rows_pad = pl.p2m_rows_pad;
  if (pl.p2m_rows_pad > 4 * 128) throw std::runtime_error("expansion order too large for the P2M kernel");
  c->p2m.upload(pl.p2m, s);
  c->p2m.mat.alloc(pl.p2m.mat_size);
  fill(c, c->p2m.mat.p, pl.p2m.mat_size, 0.);  // rows past C stay zero
  c->RP = (pl.C + 63) / 64 * 64;
  c->KP = (pl.C + 4 * dev::kKG - 1) / (4 * dev::kKG) * (4 * dev::kKG);
  // plane-wave composed M2L operators: kind 4 (10-digit rule) or 5 (6-digit); m2l_kernel 4 uses them for the FP64
  // operator too, 3 only for the FP32 one (its own tiles, below)
  const int pw_kind = c->opts.m2l_kernel >= 3 ? (c->opts.plane_wave_digits == 6 ? 5 : 4) : 0;
  const std::vector<int32_t> slot = update_operator_table(c, pl, c->opts.m2l_kernel == 4 ? pw_kind : 0);
  c->m2m.clear();
  c->l2l.clear();
  c->m2m.resize(pl.n_levels);
  c->l2l.resize(pl.n_levels);
  for (int l = 1; l < pl.n_levels; ++l) {
    c->m2m[l].upload(pl.m2m_by_level[l], s, slot);
    c->l2l[l].upload(pl.l2l_by_level[l], s, slot);
  }
  c->m2l.upload(pl.m2l, s, slot, c->opts.m2l_tile_pairs > 0 ? c->opts.m2l_tile_pairs : kTPM2L);
  c->m2l_slot = slot;
  c->m2l_min_level = pl.n_levels;  // no M2L pairs: no translations at all
  for (int32_t d : pl.m2l.dst) c->m2l_min_level = std::min(c->m2l_min_level, (int)pl.box_level[d]);
  {
    std::vector<int64_t> tptr(pl.n_boxes + 1, 0);
    for (int32_t d : pl.m2l.dst) tptr[d + 1]++;
    for (int b = 0; b < pl.n_boxes; ++b) tptr[b + 1] += tptr[b];
    std::vector<int32_t> tp(pl.m2l.dst.size());
    std::vector<int64_t> fillp(tptr.begin(), tptr.end() - 1);
    for (int64_t q = 0; q < (int64_t)pl.m2l.dst.size(); ++q) tp[fillp[pl.m2l.dst[q]]++] = (int32_t)q;
    c->m2l_tptr.upload(tptr, s);
    c->m2l_tpairs.upload(tp, s);
    std::vector<int32_t> rorder(pl.n_boxes);
    for (int b = 0; b < pl.n_boxes; ++b) rorder[b] = b;
    std::stable_sort(rorder.begin(), rorder.end(),
                     [&](int32_t x, int32_t y) { return tptr[x + 1] - tptr[x] > tptr[y + 1] - tptr[y]; });
    c->m2l_rorder.upload(rorder, s);
    // FP64 [pair][layer][C], or FP32 [pair][layer][stage_stride<float>(C)] in the same buffer
    const size_t col = std::max<size_t>(pl.C, (dev::stage_stride<float>(pl.C) + 1) / 2);
    c->m2l_stage.alloc(std::max<size_t>(1, pl.m2l.dst.size() * 2 * col));
  }
  // ---- P2M matrices on the GPU ----
  if (c->p2m.nblk > 0) {
    if ((int)pl.p2m_face_ptr.size() != c->p2m.nblk + 1 || pl.p2m_face_lc.size() != (size_t)npf * pl.p2m_faces.size())
      throw std::runtime_error("P2M source lists do not match the P2M blocks");
    const int NT = (pl.P + 1) * (pl.P + 2) / 2;
    std::vector<double> cnm(NT);
    for (int n = 0; n <= pl.P; ++n)
      for (int m = 0; m <= n; ++m) {
        double f1 = 1, f2 = 1;
        for (int k = 2; k <= n - m; ++k) f1 *= k;
        for (int k = 2; k <= n + m; ++k) f2 *= k;
        const double v = std::sqrt(f1 * f2);
        cnm[n * (n + 1) / 2 + m] = (m & 1) ? -v : v;
      }
    auto& fptr = c->sb.fptr;
    auto &pfaces = c->sb.pfaces, &plc = c->sb.plc;
    auto &pcs = c->sb.pcs, &dcnm = c->sb.dcnm;
    fptr.upload(pl.p2m_face_ptr, s);
    {
      std::vector<int32_t> pf(pl.p2m_faces.size());
      for (size_t k = 0; k < pf.size(); ++k) pf[k] = face_pos[pl.p2m_faces[k]];
      pfaces.upload(pf, s);
    }
    plc.upload(pl.p2m_face_lc, s);
    pcs.upload(pl.p2m_leaf_cs, s);
    dcnm.upload(cnm, s);
    dev::P2MBuild a;
    a.q = o.q;
    a.npf = npf;
    a.CP = pl.p2m_rows_pad;
    a.nleaves = c->p2m.nblk;
    a.face_ptr = fptr.p;
    a.faces = pfaces.p;
    a.face_lc = plc.p;
    a.leaf_cs = pcs.p;
    a.mat_ptr = c->p2m.mat_ptr.p;
    a.col_ptr = c->p2m.col_ptr.p;
    a.fX = fX.p;
    a.fcr = fcr.p;
    a.fJ = fJ.p;
    a.rule = d_dun.p;
    a.cnm = dcnm.p;
    dev::p2m_build(s, pl.P, a, c->p2m.mat.p);
    HIPCHECK(hipGetLastError());
  }
  HIPCHECK(hipStreamSynchronize(s));
  float nq_ms = 0.f;
  HIPCHECK(hipEventElapsedTime(&nq_ms, c->prep_ev[0], c->prep_ev[1]));
  const double t3 = now();
  // the near quadrature ran beside the host work of this stage: its own time from the events, the rest of the
  // stage as the upload time
  const double t_near_quad = 1e-3 * nq_ms, t_upload = std::max(0., (t3 - t1) - t_near_quad);
  // ---- rigid-mode alpha (updateGeometricProperties) ----
  evaluate(c, 2, nullptr, c->alpha.p);
  HIPCHECK(hipStreamSynchronize(s));
  const double t3a = now();
  tune_m2l_tiles(c, pl);
  HIPCHECK(hipStreamSynchronize(s));
  const double t3b = now();
  const int pc = c->opts.preconditioner;
  c->sw_host = bemfmm::SchwarzBlocks{};
  for (auto* b : {&c->sw_inv}) b->release();
  // the reference's sparse matrix (buildSparseMatrixForILU) for ILU(0) and the reference's SCHWARZ: built on the
  // host from the scaled geometry and alpha
  auto sparse_rows = [&](bemfmm::SparseRows& A) {
    std::vector<double> ah(c->n_dof);
    HIPCHECK(hipMemcpyAsync(ah.data(), c->alpha.p, c->n_dof * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    bemfmm::build_ilu_rows(c->n_nodes, pl.Xs.data(), c->n_faces, pl.fv.data(), c->face_dof_i.data(), c->n_dof,
                           c->kind_i.data(), c->point_i.data(), ah.data(), pl.scale, c->opts.ilu_neighborhood,
                           c->opts.ilu_kring, A);
  };
  if (pc == BEMFMM_PRECOND_SCHWARZ && c->opts.schwarz_variant == 1) {
    // SchwarzAdditive_CRS on the reference's blocks (GMRES_FMM.hpp:901-1013): blocks on the matrix graph, every
    // block's full inverse on the GPU (Gauss-Jordan in LDS), averaged combination
    bemfmm::SparseRows A;
    sparse_rows(A);
    bemfmm::build_schwarz_blocks_rows(A, c->opts.schwarz_core_k, c->opts.schwarz_overlap, c->opts.schwarz_max_core,
                                      c->opts.schwarz_max_block, c->sw_host);
    const bemfmm::SchwarzBlocks& B = c->sw_host;
    const int nb = B.n_blocks();
    if (B.max_block > kSchwarzMaxBlock) throw std::runtime_error("schwarz: block larger than the kernel's limit");
    std::vector<int64_t> invptr(nb + 1, 0);
    std::vector<int32_t> rows_out(nb);
    for (int b = 0; b < nb; ++b) {
      const int64_t m = B.ptr[b + 1] - B.ptr[b];
      rows_out[b] = (int32_t)m;
      invptr[b + 1] = invptr[b] + m * m;
    }
    std::vector<int64_t> occ_ptr(c->n_dof + 1, 0), occ_pos(B.idx.size());
    for (int32_t g : B.idx) occ_ptr[g + 1]++;
    for (int g = 0; g < c->n_dof; ++g) occ_ptr[g + 1] += occ_ptr[g];
    {
      std::vector<int64_t> fillp(occ_ptr.begin(), occ_ptr.end() - 1);
      for (size_t q = 0; q < B.idx.size(); ++q) occ_pos[fillp[B.idx[q]]++] = (int64_t)q;
    }
    c->ilu_ptr.upload(A.ptr, s);
    c->ilu_col.upload(A.col, s);
    c->ilu_val.upload(A.val, s);
    c->ilu_nnz = (int64_t)A.col.size();
    c->sw_ptr.upload(B.ptr, s);
    c->sw_idx.upload(B.idx, s);
    c->sw_ncore.upload(B.ncore, s);
    c->sw_invptr.upload(invptr, s);
    c->sw_rows_out.upload(rows_out, s);
    c->sw_occ_ptr.upload(occ_ptr, s);
    c->sw_occ_pos.upload(occ_pos, s);
    c->sw_y.alloc(B.idx.size());
    c->sw_inv.alloc((size_t)invptr[nb]);
    auto& failed = c->sb.failed;
    failed.alloc(nb);
    dev::schwarz_build_rows(s, schwarz_view(c), B.max_block, dev::IluRows{c->ilu_ptr.p, c->ilu_col.p, c->ilu_val.p, nullptr},
                            c->sw_rows_out.p, c->sw_inv.p, failed.p);
    HIPCHECK(hipGetLastError());
    std::vector<int32_t> fh(nb);
    HIPCHECK(hipMemcpyAsync(fh.data(), failed.p, nb * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    c->sw_failed = 0;
    for (int32_t f : fh) c->sw_failed += f;
  } else if (pc == BEMFMM_PRECOND_JACOBI || pc == BEMFMM_PRECOND_SCHWARZ) {
    auto& d_node_row = c->sb.node_row;  // every node is the target of one near row (node -> row)
    dev::NearRows nr;
    nr.kind = c->kind.p;
    nr.kappa = c->robin.p;
    nr.point = c->point.p;
    nr.node_row = d_node_row.p;
    nr.row_block = row_block.p;
    nr.row_ptr = c->near.row_ptr.p;
    nr.col_ptr = c->near.col_ptr.p;
    nr.mat_ptr = c->near.mat_ptr.p;
    nr.col_idx = c->near.col_idx.p;
    nr.ncol_dof = ncol_dof.p;
    nr.mat = c->near.mat.p;
    nr.alpha = c->alpha.p;
    if (pc == BEMFMM_PRECOND_JACOBI) {
      dev::diag_inv(s, c->n_dof, nr, c->jacobi.p);
      HIPCHECK(hipGetLastError());
    } else {
      // seeds in a spatially coherent order: the internal (Morton) DOF order, or the order of the DOFs'
      // target boxes (level-major preorder) when the caller's numbering is kept
      std::vector<int32_t> seed(c->n_dof);
      for (int i = 0; i < c->n_dof; ++i) seed[i] = i;
      if (!c->opts.dof_reorder)
        std::stable_sort(seed.begin(), seed.end(), [&](int32_t a, int32_t b) {
          return pl.target_box[c->point_i[a]] < pl.target_box[c->point_i[b]];
        });
      uint64_t key = 0xcbf29ce484222325ull;
      auto mix = [&](uint64_t v) { key = (key ^ v) * 0x100000001b3ull; };
      mix((uint64_t)c->n_dof);
      mix((uint64_t)c->n_faces);
      for (int32_t v : c->face_dof) mix((uint32_t)v);
      for (int32_t v : c->dof_point) mix((uint32_t)v);
      mix((uint64_t)c->opts.schwarz_max_core);
      mix((uint64_t)c->opts.schwarz_max_block);
      mix((uint64_t)c->opts.schwarz_overlap);
      mix((uint64_t)c->opts.dof_reorder);
      if (key == c->sw_topo_key && c->sw_caller.n_blocks() > 0) {
        // same DOF graph: the blocks of the first prepare, in this prepare's internal numbering
        c->sw_host = c->sw_caller;
        for (int32_t& g : c->sw_host.idx) g = c->inv[g];
      } else {
        bemfmm::build_schwarz_blocks(c->n_dof, c->n_faces, c->face_dof_i.data(), c->point_i.data(), seed,
                                     c->opts.schwarz_max_core, c->opts.schwarz_max_block, c->opts.schwarz_overlap,
                                     c->sw_host, pl.npf);
        c->sw_caller = c->sw_host;
        for (int32_t& g : c->sw_caller.idx) g = c->perm[g];
        c->sw_topo_key = key;
      }
      const bemfmm::SchwarzBlocks& B = c->sw_host;
      const int nb = B.n_blocks();
      if (B.max_block > kSchwarzMaxBlock) throw std::runtime_error("schwarz: block larger than the kernel's limit");
      std::vector<int64_t> invptr(nb + 1, 0);
      for (int b = 0; b < nb; ++b) invptr[b + 1] = invptr[b] + (int64_t)B.ncore[b] * (B.ptr[b + 1] - B.ptr[b]);
      c->sw_ptr.upload(B.ptr, s);
      c->sw_idx.upload(B.idx, s);
      c->sw_ncore.upload(B.ncore, s);
      c->sw_invptr.upload(invptr, s);
      c->sw_inv.alloc((size_t)invptr[nb]);
      auto& failed = c->sb.failed;
      failed.alloc(nb);
      dev::schwarz_build(s, schwarz_view(c), B.max_block, nr, c->sw_inv.p, failed.p);
      HIPCHECK(hipGetLastError());
      std::vector<int32_t> fh(nb);
      HIPCHECK(hipMemcpyAsync(fh.data(), failed.p, nb * sizeof(int32_t), hipMemcpyDeviceToHost, s));
      HIPCHECK(hipStreamSynchronize(s));
      c->sw_failed = 0;
      for (int32_t f : fh) c->sw_failed += f;
    }
  } else if (pc == BEMFMM_PRECOND_ILU0 || pc == BEMFMM_PRECOND_MILU0 || pc == BEMFMM_PRECOND_ILUT) {
    // ILU(0) / MILU(0) / ILUT factorised on the host as ILU0_CRS and ILUT_CRS do; the sweeps run on the GPU by
    // dependency level
    bemfmm::SparseRows A;
    sparse_rows(A);
    if (pc == BEMFMM_PRECOND_ILUT) {
      bemfmm::SparseRows F;
      bemfmm::ilut_factor(A, c->opts.ilut_drop_tol, c->opts.ilut_max_entries, c->opts.ilut_pivot_min, F);
      A = std::move(F);
    } else {
      bemfmm::ilu0_factor(A, 1e-20, pc == BEMFMM_PRECOND_MILU0, c->opts.milu_omega);
    }
    bemfmm::SweepLevels Lv, Uv;
    bemfmm::sweep_levels(A, Lv, Uv);
    c->ilu_nnz = (int64_t)A.col.size();
    c->ilu_ptr.upload(A.ptr, s);
    c->ilu_diag.upload(A.diag, s);
    c->ilu_col.upload(A.col, s);
    c->ilu_val.upload(A.val, s);
    c->ilu_lrows.upload(Lv.rows, s);
    c->ilu_lptr.upload(Lv.ptr, s);
    c->ilu_urows.upload(Uv.rows, s);
    c->ilu_uptr.upload(Uv.ptr, s);
    c->ilu_lptr_h = Lv.ptr;
    c->ilu_uptr_h = Uv.ptr;
    // runs of small levels in one workgroup each, large levels alone
    auto segments = [](const std::vector<int32_t>& ptr, std::vector<bemfmm_ctx::IluSegment>& seg) {
      seg.clear();
      const int nl = (int)ptr.size() - 1;
      for (int l = 0; l < nl;) {
        if (ptr[l + 1] - ptr[l] >= dev::kIluBigLevel) {
          seg.push_back({l, l + 1, true});
          ++l;
          continue;
        }
        int e = l;
        while (e < nl && ptr[e + 1] - ptr[e] < dev::kIluBigLevel) ++e;
        seg.push_back({l, e, false});
        l = e;
      }
    };
    segments(Lv.ptr, c->ilu_lseg);
    segments(Uv.ptr, c->ilu_useg);
    c->ilu_levels = Lv.n_levels() + Uv.n_levels();
  } else if (pc != BEMFMM_PRECOND_NONE) {
    throw std::runtime_error("unknown preconditioner");
  }
  HIPCHECK(hipStreamSynchronize(s));
  const double t3c = now();
  if (c->opts.precision == BEMFMM_PRECISION_FP32_REFINE) {
    c->P_in = c->opts.inner_p;  // validated above: below the plan's order, with the rotation M2L
    c->m2l_rot = (c->opts.m2l_kernel == 0 || c->opts.m2l_kernel == 2) && (c->P_in || dev::m2l_rot_compiled(pl.P));
    c->C_in = c->RP_in = c->KP_in = 0;
    if (c->m2l_rot && c->P_in) {
      // compact expansions of the inner order: M2M and L2L on the leading C_in x C_in blocks of the octant
      // operators (the packing is degree-major and M2M / L2L are triangular in the degree, so these blocks are
      // the inner order's operators), laid out [op][KP_in][RP_in] as the matrix-core tiles read them
      c->C_in = (c->P_in + 1) * (c->P_in + 1);
      c->RP_in = (c->C_in + 63) / 64 * 64;
      c->KP_in = (c->C_in + 4 * dev::kKG - 1) / (4 * dev::kKG) * (4 * dev::kKG);
      const int64_t per = (int64_t)c->KP_in * c->RP_in;
      std::vector<double> t((size_t)16 * per, 0.);
      for (int o = 0; o < 16; ++o) {
        const auto key = bemfmm::op_key(pl, o);
        if (key[0] != (o < 8 ? 0 : 1)) throw std::runtime_error("inner order: operators 0..15 are not M2M / L2L");
        const auto v = bemfmm::cached_operator(pl.P, key[0], key[1], key[2], key[3]);
        for (int k = 0; k < c->C_in; ++k)
          for (int r = 0; r < c->C_in; ++r) t[(size_t)o * per + (size_t)k * c->RP_in + r] = (*v)[(size_t)k * pl.C + r];
      }
      c->ops_in.upload(t, s);
      // per tile of m2m[l] / l2l[l] (same pairs, same tiling), the plan's operator index 0..15 = its ops_in slot
      auto tile_ops = [&](const bemfmm::PairGroups& g, const DTiles& base, DBuf<int32_t>& out) {
        std::vector<int32_t> top;
        for (size_t k = 0; k < g.op.size(); ++k)
          for (int64_t a = g.pair_ptr[k]; a < g.pair_ptr[k + 1]; a += base.tp) top.push_back(g.op[k]);
        if ((int)top.size() != base.ntiles) throw std::runtime_error("inner order: M2M / L2L tiles do not match");
        out.upload(top, s);
      };
      c->m2m_in_op.resize(pl.n_levels);
      c->l2l_in_op.resize(pl.n_levels);
      for (int l = 1; l < pl.n_levels; ++l) {
        tile_ops(pl.m2m_by_level[l], c->m2m[l], c->m2m_in_op[l]);
        tile_ops(pl.l2l_by_level[l], c->l2l[l], c->l2l_in_op[l]);
      }
    }
    // round the resident FP64 matrices once to FP32 (operators were computed in binary128 and rounded to FP64)
    {
      // the near field folded for A x: per block rows x (DOF columns + 2 Cl L2P columns, padded to 4)
      const int nb = c->near.nblk;
      std::vector<int64_t> rp;
      c->near.row_ptr.download(rp, nb + 1, s);
      std::vector<int64_t> ucp(nb + 1, 0), ump(nb, 0);
      int64_t msz = 0;
      // L2P columns per layer: the inner order's coefficients when the FP32 M2L runs below p (the locals are
      // zero above it), else all C
      const int Cl = (c->m2l_rot && c->P_in) ? c->C_in : pl.C;
      for (int b = 0; b < nb; ++b) {
        const int64_t ncu = (pl.near_ncol_dof[b] + 2 * (int64_t)Cl + 3) / 4 * 4;
        ucp[b + 1] = ucp[b] + ncu;
        ump[b] = msz;
        msz += (rp[b + 1] - rp[b]) * ncu;
      }
      if ((int64_t)c->n_dof + 2 * (int64_t)pl.n_boxes * pl.C >= (int64_t(1) << 31))
        throw std::runtime_error("FP32 near field: column index beyond 32 bits");
      c->nu_col_ptr.upload(ucp, s);
      c->nu_mat_ptr.upload(ump, s);
      c->nu_col_idx.alloc(std::max<int64_t>(1, ucp[nb]));
      c->near_mat32.alloc(std::max<int64_t>(1, msz));
      dev::fold_near(s, c->near.view(c->near_ncd.p), pl.C, Cl, c->n_dof, (int64_t)pl.n_boxes * Cl, c->near.mat.p,
                     c->kind.p, c->robin.p, c->nu_col_ptr.p, c->nu_mat_ptr.p, c->nu_col_idx.p, c->near_mat32.p);
    }
    c->p2m_mat32.alloc(c->p2m.mat.n);
    dev::to_f32(s, (int64_t)c->p2m.mat.n, c->p2m.mat.p, c->p2m_mat32.p);
    c->m2l_pw_on = c->opts.m2l_kernel == 3;
    if (c->m2l_pw_on) {
      // the FP32 operator's M2L through plane waves: extra table slots for the composed operators, own tiles
      const std::vector<int32_t> slot_pw = update_operator_table(c, pl, pw_kind);
      c->m2l_pw.upload(pl.m2l, s, slot_pw, c->m2l_tp);
    }
    if (c->m2l_rot) {
      if (c->P_in) update_rotation_table(c, pl, c->m2l_slot, c->rot_in, c->P_in);
      else update_rotation_table(c, pl, c->m2l_slot, c->rot, pl.P);
      c->m2l_rt.upload(pl.m2l, s, c->m2l_slot, kRotTP);
      c->m32.alloc((size_t)pl.n_boxes * 2 * pl.C);
      c->ops32.release();  // M2M and L2L use the FP64 table
      c->ops32_valid = 0;
    } else {
      // FP32 operator table for the dense M2L: convert the slots that are new (all after the table grew)
      const int64_t per = (int64_t)c->KP * c->RP;
      if (c->ops32.n != c->ops.n) {
        c->ops32.alloc(c->ops.n);
        c->ops32_valid = 0;
      }
      if (c->ops32_valid < c->ops_used) {
        const int64_t n32 = (int64_t)(c->ops_used - c->ops32_valid) * per;
        dev::to_f32(s, n32, c->ops.p + c->ops32_valid * per, c->ops32.p + c->ops32_valid * per);
        c->ops32_valid = c->ops_used;
      }
    }
    HIPCHECK(hipGetLastError());
  } else {
    c->near_mat32.release();
    c->nu_col_idx.release();
    c->p2m_mat32.release();
    c->ops32.release();
    c->ops32_valid = 0;
    c->m2l_rot = false;
    c->m2l_pw_on = false;
    c->P_in = 0;
  }
  tune_launch_shapes(c, pl);
  HIPCHECK(hipStreamSynchronize(s));
  const double t4 = now();
  c->prepared = true;
  bemfmm_stats& st = c->stats;
  st = bemfmm_stats{};
  st.n_nodes = c->n_nodes;
  st.n_faces = c->n_faces;
  st.n_dof = c->n_dof;
  st.n_levels = pl.n_levels;
  st.n_boxes = pl.n_boxes;
  st.n_leaves = (int)pl.n_leaves;
  st.n_m2l_pairs = pl.n_m2l_pairs;
  st.near_nnz = pl.near_nnz;
  st.n_m2l_ops = (int)pl.m2l.op.size();
  st.L = pl.Lscale;
  st.device_bytes = (double)(c->near.bytes() + c->p2m.bytes() + c->m2l.bytes() + c->ops.bytes() + c->buf0.bytes() +
                             c->buf1.bytes() + c->M0.bytes() + c->M1.bytes() + c->m2l_stage.bytes() +
                             c->near_mat32.bytes() + c->nu_col_idx.bytes() + c->p2m_mat32.bytes() + c->ops32.bytes() + c->sw_inv.bytes() +
                             c->m2l_pw.bytes());
  st.t_tree = pl.t_tree;
  st.t_lists = pl.t_lists;
  st.t_near_struct = pl.t_near_struct;
  st.t_p2m = pl.t_p2m;
  st.t_ops = pl.t_ops;
  st.t_upload = t_upload;
  st.t_near_quad = t_near_quad;
  st.t_alpha = t3a - t3;
  st.t_tune = t3b - t3a;
  st.t_precond = t3c - t3b;
  st.t_fp32 = t4 - t3c;
  for (int k = 0; k < 2; ++k) {
    const bool used = k == 0 || c->opts.precision == BEMFMM_PRECISION_FP32_REFINE;
    st.near_rows_per_wave[k] = used ? kRowsPerWave[c->gemv_var[k]] : 0;
    st.p2m_loads[k] = used ? kP2MLoads[c->p2m_var[k]] : 0;
  }
  st.n_precond_blocks = c->sw_host.n_blocks();
  st.precond_fallback_blocks = c->sw_failed;
  st.precond_avg_block = c->sw_host.n_blocks() ? (double)c->sw_host.idx.size() / c->sw_host.n_blocks() : 0.;
  st.n_ops_uploaded = c->ops_uploaded;
  st.m2l_tile_pairs = c->m2l_tp;
  st.m2l_rotation = c->m2l_rot ? 1 : 0;
  st.inner_p = c->P_in;
  st.m2l_plane_wave = c->opts.m2l_kernel == 4 ? 2 : c->m2l_pw_on ? 1 : 0;
  const bool ilu_kind = pc == BEMFMM_PRECOND_ILU0 || pc == BEMFMM_PRECOND_MILU0 || pc == BEMFMM_PRECOND_ILUT;
  st.precond_nnz = (ilu_kind || (pc == BEMFMM_PRECOND_SCHWARZ && c->opts.schwarz_variant == 1)) ? c->ilu_nnz : 0;
  st.precond_levels = ilu_kind ? c->ilu_levels : 0;
  st.tree_on_gpu = (c->opts.tree_on_gpu && o.max_level <= 11) ? 1 : 0;
  (void)t0;
}

void ensure_basis(bemfmm_ctx* c, int64_t need) {
  if (need <= c->v_cap) return;
  int64_t cap = std::max<int64_t>(need, std::max<int64_t>(64, 2 * c->v_cap));
  DBuf<double> nv;
  nv.alloc((size_t)cap * c->n_dof);
  if (c->v_cap > 0)
    HIPCHECK(hipMemcpyAsync(nv.p, c->V.p, (size_t)c->v_cap * c->n_dof * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  c->V.swap(nv);
  c->v_cap = cap;
  c->partial.alloc((size_t)kDotBlocks * (cap + 1));
  // [h1 (k) | h2 (k) | w.w | projected coefficients (k)] of one Arnoldi step (k <= cap), or the k update
  // coefficients
  c->hbuf.alloc(3 * (size_t)cap + 3);
  if (c->hres) (void)pinned_free(c->hres);
  c->hres = nullptr;
  HIPCHECK(pinned_malloc((void**)&c->hres, (2 * (size_t)cap + 2) * sizeof(double)));
}

// hbuf[off, off + k) = V[0..k) . w on the device (no host copy)
void dots_dev(bemfmm_ctx* c, int k, const double* Vp, const double* w, size_t off) {
  const int64_t n = c->n_dof;
  // the reduction buffers are sized by ensure_basis for up to v_cap + 1 products
  if (k < 1 || k > c->v_cap + 1 || !c->partial.p || !c->hbuf.p || (size_t)k * kDotBlocks > c->partial.n ||
      off + (size_t)k > c->hbuf.n)
    throw HipError("internal: dot-product buffers not sized for " + std::to_string(k) + " products", BEMFMM_E_STATE);
  dev::dots(c->stream, k, Vp, n, w, n, c->partial.p);
  dev::reduce_partials(c->stream, c->partial.p, k, c->hbuf.p + off);
}
// h[0..k) = V[0..k) . w  (device), copied to host
void dots(bemfmm_ctx* c, int k, const double* Vp, const double* w, std::vector<double>& h) {
  dots_dev(c, k, Vp, w, 0);
  h.resize(k);
  HIPCHECK(hipMemcpyAsync(h.data(), c->hbuf.p, k * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
}
double norm2(bemfmm_ctx* c, const double* w) {
  std::vector<double> h;
  dots(c, 1, w, w, h);
  return std::sqrt(h[0]);
}

// out = M^-1 in for the context's preconditioner (left preconditioning, return_A_dot_v and the
// preconditioned RHS, BEM_solveBVP_GMRES_FMM.hpp:2124-2171). in and out may alias except for Schwarz.
void precond_apply(bemfmm_ctx* c, const double* in, double* out) {
  const int64_t n = c->n_dof;
  switch (c->opts.preconditioner) {
  case BEMFMM_PRECOND_JACOBI:
    if (in != out) dev::mul_to(c->stream, n, c->jacobi.p, in, out);
    else dev::mul(c->stream, n, c->jacobi.p, out);
    break;
  case BEMFMM_PRECOND_SCHWARZ:
    if (in == out) throw std::runtime_error("internal: Schwarz apply needs distinct buffers");
    if (c->opts.schwarz_variant == 1)
      dev::schwarz_apply_averaged(c->stream, schwarz_view(c), (int)n, c->sw_occ_ptr.p, c->sw_occ_pos.p, c->sw_inv.p, in,
                                  c->sw_y.p, out);
    else
      dev::schwarz_apply(c->stream, schwarz_view(c), c->sw_inv.p, in, out);
    break;
  case BEMFMM_PRECOND_ILU0:
  case BEMFMM_PRECOND_MILU0:
  case BEMFMM_PRECOND_ILUT: {
    if (in != out) HIPCHECK(hipMemcpyAsync(out, in, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    dev::IluRows A{c->ilu_ptr.p, c->ilu_col.p, c->ilu_val.p, c->ilu_diag.p};
    for (const auto& g : c->ilu_lseg) {
      if (g.single) dev::ilu_level(c->stream, A, c->ilu_lrows.p + c->ilu_lptr_h[g.l0],
                                   c->ilu_lptr_h[g.l0 + 1] - c->ilu_lptr_h[g.l0], false, out);
      else dev::ilu_run(c->stream, A, c->ilu_lrows.p, c->ilu_lptr.p, g.l0, g.l1, false, out);
    }
    for (const auto& g : c->ilu_useg) {
      if (g.single) dev::ilu_level(c->stream, A, c->ilu_urows.p + c->ilu_uptr_h[g.l0],
                                   c->ilu_uptr_h[g.l0 + 1] - c->ilu_uptr_h[g.l0], true, out);
      else dev::ilu_run(c->stream, A, c->ilu_urows.p, c->ilu_uptr.p, g.l0, g.l1, true, out);
    }
    break;
  }
  default:
    if (in != out) HIPCHECK(hipMemcpyAsync(out, in, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  }
}

// w = M^-1 A v (return_A_dot_v); lo: FP32 operator. With the graph replay the preconditioner reads the
// graph's output buffer (no copy out; gout is next written by the next replay on the same stream).
void apply_op(bemfmm_ctx* c, const double* v, double* w, bool lo) {
  double* y = graph_mode(c) ? c->gout.p : c->opts.preconditioner == BEMFMM_PRECOND_SCHWARZ ? c->ptmp.p : w;
  evaluate(c, 0, v, y, lo);
  precond_apply(c, y, w);
}

// w *= a for a new basis vector; with the graph replay it is also written to the graph's input buffer,
// so the next evaluate(w) starts without a copy
void scale_basis(bemfmm_ctx* c, double a, double* w) {
  const int64_t n = c->n_dof;
  if (graph_mode(c)) {
    dev::scale_dup(c->stream, n, a, w, c->gin.p);
    c->gin_holds = w;
  } else {
    dev::axpby(c->stream, n, a, w, 0.0, w);
  }
}

// One GMRES run on  M^-1 A x = rhs  (rhs: device, already preconditioned), x (device) updated
// in place. confirm = true follows run_gmres (BEM_solveBVP_GMRES_FMM.hpp:2275-2355): restart
// cycles, an estimate <= 10 tol triggers a true-residual check, up to 5 saturated checks.
// confirm = false (inner solve of the refinement loop) stops when the estimate reaches tol.
// Arnoldi + Givens as basic_linear_systems.hpp:3137-3451, with classical Gram-Schmidt applied
// twice on the device. Returns the last (true or estimated) residual; iters accumulates.
struct GmresRun {
  const double* rhs;
  double* x;
  double tol;
  int max_iter;  // iteration budget including what `iters` holds on entry
  int restart;
  bool lo;       // FP32 matvecs
  bool confirm;
  bool x_zero;   // x is zero on entry: r0 = rhs without a matvec
};

double gmres_run(bemfmm_ctx* c, const GmresRun& R, int& iters) {
  const int64_t n = c->n_dof;
  hipStream_t s = c->stream;
  auto Av = [&](const double* v, double* w) { apply_op(c, v, w, R.lo); };
  double cur_err = 1e300;
  const int restart_max = std::max(1, R.restart);
  bool first = true;
  while (iters < R.max_iter && cur_err > R.tol) {
    const int remaining = R.max_iter - iters;
    const int k_cycle_max = std::min(restart_max, remaining);
    ensure_basis(c, 8);
    // r0 = rhs - M^-1 A x
    if (first && R.x_zero) {
      HIPCHECK(hipMemcpyAsync(c->V.p, R.rhs, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    } else {
      Av(R.x, c->tmpv.p);
      dev::sub(s, n, R.rhs, c->tmpv.p, c->V.p);
    }
    first = false;
    const double beta = norm2(c, c->V.p);
    HIPCHECK(hipMemcpyAsync(c->xbase.p, R.x, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    std::vector<double> cs, sn, g(1, beta);
    std::vector<std::vector<double>> Rm;
    if (!std::isfinite(beta)) throw HipError("GMRES breakdown: non-finite initial residual", BEMFMM_E_NONFINITE);
    if (beta == 0.) {
      cur_err = 0.;
      break;
    }
    scale_basis(c, 1.0 / beta, c->V.p);
    int nb = 1;  // basis vectors in V
    auto add_basis = [&]() {
      ensure_basis(c, nb + 1);
      double* w = c->V.p + (int64_t)nb * n;
      Av(c->V.p + (int64_t)(nb - 1) * n, w);
      // both Gram-Schmidt passes and the norm run on the device into hbuf = [h1 | h2 | w.w]; the
      // column comes to the host once per iteration
      dots_dev(c, nb, c->V.p, w, 0);
      dev::vsum(s, c->V.p, n, nb, c->hbuf.p, -1.0, w, n);
      dots_dev(c, nb, c->V.p, w, nb);
      dev::vsum(s, c->V.p, n, nb, c->hbuf.p + nb, -1.0, w, n);
      dots_dev(c, 1, w, w, 2 * (size_t)nb);
      HIPCHECK(hipMemcpyAsync(c->hres, c->hbuf.p, (2 * (size_t)nb + 1) * sizeof(double), hipMemcpyDeviceToHost, s));
      HIPCHECK(hipStreamSynchronize(s));
      std::vector<double> col(nb + 1);
      for (int i = 0; i < nb; ++i) col[i] = c->hres[i] + c->hres[nb + i];
      const double hn = std::sqrt(c->hres[2 * nb]);
      col[nb] = hn;
      scale_basis(c, hn > 0. ? 1.0 / hn : 0.0, w);
      ++nb;
      // Givens update of the new column
      const size_t cc = cs.size();
      for (size_t i = 0; i < cc; ++i) {
        const double tmp = std::fma(cs[i], col[i], sn[i] * col[i + 1]);
        col[i + 1] = std::fma(cs[i], col[i + 1], -sn[i] * col[i]);
        col[i] = tmp;
      }
      double cg = 1.0, sg = 0.0;
      const double r = std::hypot(col[cc], col[cc + 1]);
      if (r != 0.0) {
        cg = col[cc] / r;
        sg = col[cc + 1] / r;
      }
      const double tmp = std::fma(cg, col[cc], sg * col[cc + 1]);
      col[cc + 1] = std::fma(cg, col[cc + 1], -sg * col[cc]);
      col[cc] = tmp;
      cs.push_back(cg);
      sn.push_back(sg);
      g.push_back(0.);
      const double gt = std::fma(cg, g[cc], sg * g[cc + 1]);
      g[cc + 1] = std::fma(cg, g[cc + 1], -sg * g[cc]);
      g[cc] = gt;
      Rm.push_back(col);
    };
    auto update_solution = [&]() {
      const size_t k = cs.size();
      std::vector<double> y(k, 0.);
      for (int i = (int)k - 1; i >= 0; --i) {
        double sum = g[i];
        for (size_t j = i + 1; j < k; ++j) sum -= Rm[j][i] * y[j];
        y[i] = sum / Rm[i][i];
      }
      HIPCHECK(hipMemcpyAsync(R.x, c->xbase.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
      HIPCHECK(hipMemcpyAsync(c->hbuf.p, y.data(), k * sizeof(double), hipMemcpyHostToDevice, s));
      dev::vsum(s, c->V.p, n, (int)k, c->hbuf.p, 1.0, R.x, n);
      HIPCHECK(hipStreamSynchronize(s));
    };
    auto true_residual = [&]() {
      Av(R.x, c->tmpv.p);
      dev::sub(s, n, c->tmpv.p, R.rhs, c->tmpv.p);
      return norm2(c, c->tmpv.p);
    };
    add_basis();  // gmres(A, b, x, 1)
    iters += 1;
    c->history.push_back(std::abs(g.back()));
    if (!R.confirm && std::abs(g.back()) <= R.tol) {
      update_solution();
      cur_err = std::abs(g.back());
      break;
    }
    int sat = 0;
    while (true) {
      add_basis();
      ++iters;
      const double est = std::abs(g.back());
      c->history.push_back(est);
      if (!std::isfinite(est)) throw HipError("GMRES breakdown: non-finite residual estimate", BEMFMM_E_NONFINITE);
      if (!R.confirm) {
        if (est <= R.tol || (int)cs.size() >= k_cycle_max) {
          update_solution();
          cur_err = est;
          break;
        }
        continue;
      }
      if (est <= R.tol * 10.) {
        update_solution();
        cur_err = true_residual();
        if (cur_err <= R.tol) break;
        if (sat++ >= 5) break;
      }
      if (iters >= R.max_iter && (int)cs.size() >= k_cycle_max) {
        update_solution();
        cur_err = (est <= R.tol * 10.) ? true_residual() : est;
        break;
      }
    }
    if (!R.confirm) break;
  }
  return cur_err;
}

// Least-squares problem min |g - H y| of a Krylov solve whose columns may have several entries below the
// diagonal, by Givens rotations: each new column gets the rotations so far, then its entries below the diagonal
// are zeroed bottom-up (rows (i-1, i)), and the same rotations go to g. With one entry below the diagonal (an
// Arnoldi column) this is the Givens update of basic_linear_systems.hpp:3137-3451.
struct GivensLsq {
  struct Rot {
    int i;
    double c, s;
  };
  std::vector<Rot> rot;
  std::vector<std::vector<double>> R;  // rotated columns (upper triangular part used)
  std::vector<double> g;               // rotated right-hand side
  static void apply(const Rot& r, std::vector<double>& v) {
    if (r.i + 1 >= (int)v.size()) return;
    const double a = v[r.i], b = v[r.i + 1];
    v[r.i] = std::fma(r.c, a, r.s * b);
    v[r.i + 1] = std::fma(r.c, b, -r.s * a);
  }
  void reset(const std::vector<double>& rhs) {
    rot.clear();
    R.clear();
    g = rhs;
  }
  void add_column(std::vector<double> col) {
    if (col.size() > g.size()) g.resize(col.size(), 0.);
    for (const Rot& r : rot) apply(r, col);
    const int k = (int)R.size();
    for (int i = (int)col.size() - 1; i > k; --i) {
      if (col[i] == 0.) continue;
      const double h = std::hypot(col[i - 1], col[i]);
      const Rot r{i - 1, col[i - 1] / h, col[i] / h};
      apply(r, col);
      col[i] = 0.;
      apply(r, g);
      rot.push_back(r);
    }
    R.push_back(std::move(col));
  }
  // |g - H y| at the minimiser: the rotated right-hand side below the triangle
  double residual() const {
    double s = 0.;
    for (size_t i = R.size(); i < g.size(); ++i) s += g[i] * g[i];
    return std::sqrt(s);
  }
  std::vector<double> solve() const {
    const int k = (int)R.size();
    std::vector<double> y(k, 0.);
    for (int i = k - 1; i >= 0; --i) {
      double sum = g[i];
      for (int j = i + 1; j < k; ++j) sum -= R[j][i] * y[j];
      y[i] = sum / R[i][i];
    }
    return y;
  }
};

// Krylov space of the first inner solve of a refinement loop, kept for the second: orthonormal basis
// W_0..W_{nw-1} (the context's basis buffer) and the Arnoldi columns H[j] with M^-1 A_32 W_j = W H[j].
// Every inner solve has the same operator M^-1 A_32. After a plain restart, GMRES has to resolve again the
// slowly converging directions that the first solve already found; GCRO (de Sturler 1996; Parks et al.
// 2006) keeps C = range(M^-1 A_32 W_{0..k-1}) and runs the second solve's Arnoldi on (I - C C^T) M^-1 A_32.
struct RecycledKrylov {
  int nw = 0;
  std::vector<std::vector<double>> H;
};

// Thin QR of the nw x k matrix with columns H (modified Gram-Schmidt, twice): Q (k columns of length nw),
// R (k columns, upper triangular). Returns false if a column is dependent on the earlier ones.
bool small_qr(const std::vector<std::vector<double>>& H, int nw, std::vector<std::vector<double>>& Q,
              std::vector<std::vector<double>>& R) {
  const int k = (int)H.size();
  Q.assign(k, std::vector<double>(nw, 0.));
  R.assign(k, std::vector<double>(k, 0.));
  for (int j = 0; j < k; ++j) {
    std::vector<double> v(nw, 0.);
    for (size_t i = 0; i < H[j].size() && (int)i < nw; ++i) v[i] = H[j][i];
    double n0 = 0.;
    for (double x : v) n0 += x * x;
    for (int pass = 0; pass < 2; ++pass)
      for (int i = 0; i < j; ++i) {
        double h = 0.;
        for (int t = 0; t < nw; ++t) h += Q[i][t] * v[t];
        for (int t = 0; t < nw; ++t) v[t] -= h * Q[i][t];
        R[j][i] += h;
      }
    double nn = 0.;
    for (double x : v) nn += x * x;
    nn = std::sqrt(nn);
    if (!(nn > 1e-12 * std::sqrt(n0))) return false;
    R[j][j] = nn;
    for (int t = 0; t < nw; ++t) Q[j][t] = v[t] / nn;
  }
  return true;
}

// One inner solve of the refinement loop: d (device, overwritten) ~ argmin |r - M^-1 A_32 d|, until the
// estimate is <= tol or `budget` matvecs are spent. Without kept state this is GMRES from r and its Krylov
// space is kept in K; with kept state (the first solve's), the search space is the kept directions plus the
// Krylov space of (I - C C^T) M^-1 A_32 from (I - C C^T) r, new vectors orthogonalised against C (held as
// combinations Q_C of the kept basis) and against each other, both twice (classical Gram-Schmidt); K is
// emptied afterwards (a third step restarts). Returns the estimate; iters counts the matvecs.
double refine_inner(bemfmm_ctx* c, RecycledKrylov& K, const double* r, double* d, double tol, int budget,
                    int max_basis, int& iters) {
  const int64_t n = c->n_dof;
  hipStream_t s = c->stream;
  auto basis = [&](int i) { return c->V.p + (int64_t)i * n; };
  std::vector<std::vector<double>> QC, RC;
  const bool recycle = !K.H.empty() && K.nw + 2 < max_basis && small_qr(K.H, K.nw, QC, RC);
  const int nk = recycle ? K.nw : 0;      // kept physical vectors W_0..W_{nk-1}
  const int kc = recycle ? (int)K.H.size() : 0;  // kept directions = columns of C
  // projection of the physical vector at position p onto C (coefficients over W in `pw`, over the new
  // vectors unchanged): the first nk dot products a become Q_C Q_C^T a; returns Q_C^T a
  auto project_c = [&](std::vector<double>& coef) {
    std::vector<double> cc(kc, 0.);
    for (int j = 0; j < kc; ++j)
      for (int i = 0; i < nk; ++i) cc[j] += QC[j][i] * coef[i];
    for (int i = 0; i < nk; ++i) {
      double v = 0.;
      for (int j = 0; j < kc; ++j) v += QC[j][i] * cc[j];
      coef[i] = v;
    }
    return cc;
  };
  // w <- w minus its components along C and the new vectors [nk, nb), twice; returns the coefficients
  // (over C, then the new vectors) and |w|^2 in `norm2w`. Both passes and the norm run on the device
  // (hbuf = [dots of pass 1 | dots of pass 2 | w.w | projected coefficients]; the first nk dots are projected
  // by Q_C Q_C^T there); the raw dots come to the host once, for the coefficients over C.
  auto orthogonalise = [&](double* w, int nb, double& norm2w) {
    std::vector<double> total(kc + nb - nk, 0.);
    const size_t pb = 2 * (size_t)nb + 1;
    for (int pass = 0; pass < 2; ++pass) {
      dots_dev(c, nb, c->V.p, w, (size_t)pass * nb);
      dev::proj_coef(s, nk, nb, c->gcro_proj.p, c->hbuf.p + (size_t)pass * nb, c->hbuf.p + pb);
      dev::vsum(s, c->V.p, n, nb, c->hbuf.p + pb, -1.0, w, n);
    }
    dots_dev(c, 1, w, w, 2 * (size_t)nb);
    HIPCHECK(hipMemcpyAsync(c->hres, c->hbuf.p, pb * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<double> coef(c->hres + (size_t)pass * nb, c->hres + (size_t)(pass + 1) * nb);
      const std::vector<double> cc = project_c(coef);
      for (int j = 0; j < kc; ++j) total[j] += cc[j];
      for (int i = nk; i < nb; ++i) total[kc + i - nk] += coef[i];
    }
    norm2w = c->hres[2 * (size_t)nb];
    return total;
  };
  std::vector<double> proj_h;  // kept alive until the end of the solve (its upload is asynchronous)
  if (recycle) {
    proj_h.assign((size_t)nk * nk, 0.);
    for (int i = 0; i < nk; ++i)
      for (int i2 = 0; i2 < nk; ++i2) {
        double v = 0.;
        for (int j = 0; j < kc; ++j) v += QC[j][i] * QC[j][i2];
        proj_h[(size_t)i * nk + i2] = v;
      }
    if (c->gcro_proj.n < proj_h.size()) c->gcro_proj.alloc(proj_h.size());
    HIPCHECK(hipMemcpyAsync(c->gcro_proj.p, proj_h.data(), proj_h.size() * sizeof(double), hipMemcpyHostToDevice, s));
  }
  if (!recycle) K = RecycledKrylov{};
  // first new vector: r without its components along C
  ensure_basis(c, nk + 2);
  double* w = basis(nk);
  HIPCHECK(hipMemcpyAsync(w, r, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  std::vector<double> g;
  double zn2 = 0.;
  if (nk > 0) {
    g = orthogonalise(w, nk, zn2);  // C^T r
  } else {
    std::vector<double> h;
    dots(c, 1, w, w, h);
    zn2 = h[0];
  }
  const double zn = std::sqrt(zn2);
  if (!std::isfinite(zn)) throw HipError("refinement: non-finite residual", BEMFMM_E_NONFINITE);
  g.push_back(zn);
  GivensLsq ls;
  ls.reset(g);
  for (int j = 0; j < kc; ++j) ls.add_column(RC[j]);  // images of the kept directions, in C coordinates
  int nb = nk;
  if (zn > 0.) {
    scale_basis(c, 1.0 / zn, w);
    nb = nk + 1;
  }
  double est = ls.residual();
  int used = 0;
  bool more = zn > 0.;
  std::vector<std::vector<double>> Hnew;
  while (est > tol && used < budget && more && nb < max_basis) {
    ensure_basis(c, nb + 1);
    w = basis(nb);
    apply_op(c, basis(nb - 1), w, true);
    double hn2 = 0.;
    std::vector<double> col;
    if (nk > 0) {
      col = orthogonalise(w, nb, hn2);
    } else {
      // plain Arnoldi: both Gram-Schmidt passes and the norm on the device, one copy to the host
      dots_dev(c, nb, c->V.p, w, 0);
      dev::vsum(s, c->V.p, n, nb, c->hbuf.p, -1.0, w, n);
      dots_dev(c, nb, c->V.p, w, nb);
      dev::vsum(s, c->V.p, n, nb, c->hbuf.p + nb, -1.0, w, n);
      dots_dev(c, 1, w, w, 2 * (size_t)nb);
      HIPCHECK(hipMemcpyAsync(c->hres, c->hbuf.p, (2 * (size_t)nb + 1) * sizeof(double), hipMemcpyDeviceToHost, s));
      HIPCHECK(hipStreamSynchronize(s));
      col.resize(nb);
      for (int i = 0; i < nb; ++i) col[i] = c->hres[i] + c->hres[nb + i];
      hn2 = c->hres[2 * nb];
    }
    const double hn = std::sqrt(hn2);
    col.push_back(hn);
    scale_basis(c, hn > 0. ? 1.0 / hn : 0.0, w);
    more = hn > 0.;  // an exact breakdown leaves nothing new to expand
    ++nb;
    if (nk == 0) K.H.push_back(col);
    ls.add_column(std::move(col));
    ++used;
    ++iters;
    est = ls.residual();
    c->history.push_back(est);
    if (!std::isfinite(est)) throw HipError("GMRES breakdown: non-finite residual estimate", BEMFMM_E_NONFINITE);
  }
  // d = W_{0..kc-1} y_kept + new vectors y_new (the kept directions are the kept basis vectors themselves)
  const std::vector<double> y = ls.solve();
  std::vector<double> coef(nb, 0.);
  for (int j = 0; j < kc; ++j) coef[j] = y[j];
  for (size_t j = kc; j < y.size(); ++j) coef[nk + (j - kc)] = y[j];
  fill(c, d, n, 0.);
  if (!y.empty()) {
    HIPCHECK(hipMemcpyAsync(c->hbuf.p, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice, s));
    dev::vsum(s, c->V.p, n, nb, c->hbuf.p, 1.0, d, n);
  }
  HIPCHECK(hipStreamSynchronize(s));
  if (nk == 0) K.nw = nb;
  else K = RecycledKrylov{};
  return est;
}

// FP64 preconditioned residual r = M^-1 (b - A x) into out; returns |r|_2.
double fp64_residual(bemfmm_ctx* c, const double* x, double* out) {
  const int64_t n = c->n_dof;
  apply_op(c, x, out, false);
  dev::sub(c->stream, n, c->bvec.p, out, out);
  return norm2(c, out);
}

// Solve (solveSystemGMRES, BEM_solveBVP_GMRES_FMM.hpp:2007-2355).
// FP64: one run_gmres with FP64 matvecs.
// FP32_REFINE: iterative refinement. Each step takes the FP64 residual r = M^-1 (b - A x), solves
// M^-1 A_32 d = r with FP32 matvecs to |r| * refine_eta, and sets x += d; it stops when |r| <= tol
// (the same FP64 criterion as the FP64 path). If a step fails to halve |r|, the remaining budget
// goes to an FP64 run_gmres from the current x, so the answer always meets the FP64 criterion.
int do_solve(bemfmm_ctx* c, const double* x0, double* xout, double tol, int max_iter, int restart, bemfmm_result* res) {
  if (!c->prepared) throw std::runtime_error("bemfmm_solve: call bemfmm_prepare first");
  const double ts = now();
  const int64_t n = c->n_dof;
  hipStream_t s = c->stream;
  const int n_mv0 = c->n_matvec;
  ensure_basis(c, 8);  // Krylov basis and the dot-product buffers (norms are taken before the first cycle)
  evaluate(c, 1, nullptr, c->ptmp.p);  // b, then M^-1 b (left preconditioning, GMRES_FMM.hpp:2157-2171)
  precond_apply(c, c->ptmp.p, c->bvec.p);
  std::vector<double> x0i(n);
  host_to_internal(c, x0, x0i.data());
  // the unknowns of Dirichlet DOFs (phi_n) are solved for in scaled form, L phi_n, as the result is
  // returned divided by L below: the initial guess is scaled the same way
  const double L0 = c->plan->Lscale;
  for (int64_t i = 0; i < n; ++i)
    if (c->kind_i[i] == 0) x0i[i] *= L0;
  HIPCHECK(hipMemcpyAsync(c->xsol.p, x0i.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
  c->history.clear();
  c->refine_history.clear();
  int iters = 0, steps = 0;
  double cur_err = 0.;
  // a zero initial guess has r0 = M^-1 b: no matvec for the first residual
  const bool x0_zero = std::all_of(x0i.begin(), x0i.end(), [](double v) { return v == 0.; });
  if (c->opts.precision == BEMFMM_PRECISION_FP32_REFINE) {
    const double eta = (c->opts.refine_eta > 0. && c->opts.refine_eta < 1.) ? c->opts.refine_eta : 3e-7;
    double prev = 1e300;
    bool fallback = false;
    bool x_zero = x0_zero;
    RecycledKrylov K;
    while (true) {
      if (x_zero) {
        HIPCHECK(hipMemcpyAsync(c->rvec.p, c->bvec.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        cur_err = norm2(c, c->rvec.p);
        x_zero = false;
      } else {
        cur_err = fp64_residual(c, c->xsol.p, c->rvec.p);
      }
      c->refine_history.push_back(cur_err);
      if (!std::isfinite(cur_err)) throw HipError("non-finite residual in refinement", BEMFMM_E_NONFINITE);
      if (cur_err <= tol || iters >= max_iter) break;
      if (cur_err > 0.5 * prev) {
        fallback = true;
        break;
      }
      prev = cur_err;
      // the last step (eta |r| below the tolerance) only has to reach the tolerance: its inner solve stops at
      // 0.8 tol (the FP64 residual after it lands below the inner estimate; the check above still decides)
      const double inner_tol = std::max(eta * cur_err, 0.8 * tol);
      if (c->opts.refine_recycle) {
        // the second step reuses the first step's Krylov space (GCRO); the basis holds at most restart + 1 vectors
        const int max_basis = std::max(restart, 1) + 1;
        refine_inner(c, K, c->rvec.p, c->dvec.p, inner_tol, max_iter - iters, max_basis, iters);
      } else {
        fill(c, c->dvec.p, n, 0.);
        GmresRun in{c->rvec.p, c->dvec.p, inner_tol, max_iter, restart, true, false, true};
        gmres_run(c, in, iters);
      }
      dev::axpby(s, n, 1.0, c->dvec.p, 1.0, c->xsol.p);
      ++steps;
    }
    if (fallback && iters < max_iter) {
      GmresRun r64{c->bvec.p, c->xsol.p, tol, max_iter, restart, false, true, false};
      cur_err = gmres_run(c, r64, iters);
    }
  } else {
    GmresRun r64{c->bvec.p, c->xsol.p, tol, max_iter, restart, false, true, x0_zero};
    cur_err = gmres_run(c, r64, iters);
  }
  std::vector<double> x(n);
  HIPCHECK(hipMemcpyAsync(x.data(), c->xsol.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
  const double L = c->plan->Lscale;
  for (int64_t i = 0; i < n; ++i) xout[c->perm[i]] = (c->kind_i[i] == 0) ? x[i] / L : x[i];
  if (res) {
    res->iters = iters;
    res->n_matvec = c->n_matvec - n_mv0;
    res->converged = cur_err <= tol ? 1 : 0;
    res->residual = cur_err;
    res->time_solve = now() - ts;
    res->refinement_steps = steps;
  }
  for (int64_t i = 0; i < n; ++i)
    if (!std::isfinite(xout[i])) throw HipError("non-finite solution", BEMFMM_E_NONFINITE);
  return cur_err <= tol ? BEMFMM_OK : BEMFMM_E_NOT_CONVERGED;
}

// y = A x on host vectors in caller order
void host_matvec(bemfmm_ctx* c, const double* x, double* y, bool lo) {
  const int n = c->n_dof;
  std::vector<double> xi(n), yi(n);
  host_to_internal(c, x, xi.data());
  HIPCHECK(hipMemcpyAsync(c->xin.p, xi.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  evaluate(c, 0, c->xin.p, c->yout.p, lo);
  HIPCHECK(hipMemcpyAsync(yi.data(), c->yout.p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  host_to_caller(c, yi.data(), y);
}

// y = A x on device vectors in caller order (permuted through the context's buffers unless the order is the caller's)
void device_matvec(bemfmm_ctx* c, const double* x, double* y, bool lo) {
  const int n = c->n_dof;
  if (c->perm_identity) {
    evaluate(c, 0, x, y, lo);
    return;
  }
  dev::perm_gather(c->stream, n, c->d_perm.p, x, c->xin.p);
  evaluate(c, 0, c->xin.p, c->yout.p, lo);
  dev::perm_scatter(c->stream, n, c->d_perm.p, c->yout.p, y);
  HIPCHECK(hipGetLastError());
}

// device vector (internal order) -> host vector (caller order)
void download_caller(bemfmm_ctx* c, const double* d, double* out) {
  std::vector<double> v(c->n_dof);
  HIPCHECK(hipMemcpyAsync(v.data(), d, c->n_dof * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  host_to_caller(c, v.data(), out);
}

}  // namespace

extern "C" {

int bemfmm_is_available(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n > 0 ? 1 : 0;
}

const char* bemfmm_device_name(void) {
  static char name[600] = {0};
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, 0) != hipSuccess) return "unavailable";
  std::snprintf(name, sizeof(name), "%s (%s)", p.name, p.gcnArchName);
  return name;
}

void bemfmm_default_opts(bemfmm_opts* o) {
  if (!o) return;
  o->p = 10;
  o->theta = 0.25;
  o->leaf_max = 50;
  o->max_level = 7;
  o->q = 6;
  o->precision = BEMFMM_PRECISION_FP64;
  o->preconditioner = BEMFMM_PRECOND_NONE;
  o->refine_eta = 3e-7;
  o->schwarz_max_core = 64;   // main_time_domain.cpp:47
  o->schwarz_max_block = 128; // main_time_domain.cpp:48
  o->schwarz_overlap = 1;     // main_time_domain.cpp:46
  o->m2l_tile_pairs = 0;
  o->dof_reorder = 1;
  o->use_graphs = 0;
  o->m2l_kernel = 0;
  o->overlap_near = 1;
  o->refine_recycle = 1;
  o->ilu_neighborhood = 0;  // BUCKETS (main_time_domain.cpp:39)
  o->ilu_kring = 1;         // main_time_domain.cpp:40
  o->schwarz_variant = 0;
  o->schwarz_core_k = 1;    // main_time_domain.cpp:45
  o->milu_omega = 1.0;        // main_time_domain.cpp:41
  o->ilut_drop_tol = 1e-3;    // main_time_domain.cpp:42
  o->ilut_max_entries = 50;   // main_time_domain.cpp:43
  o->ilut_pivot_min = 1e-12;  // main_time_domain.cpp:44
  o->plane_wave_digits = 10;
  o->tree_on_gpu = 1;
  o->element_order = 1;
}

bemfmm_ctx* bemfmm_ctx_create(int device, const bemfmm_opts* opts) {
  auto* c = new bemfmm_ctx();
  c->device = device;
  if (opts) c->opts = *opts;
  else bemfmm_default_opts(&c->opts);
  const int rc = guarded(c, [&]() {
    int n = 0;
    HIPCHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) throw HipError("invalid device index", BEMFMM_E_DEVICE);
    HIPCHECK(hipSetDevice(device));
    HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHECK(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
    HIPCHECK(hipStreamCreateWithFlags(&c->gstream, hipStreamNonBlocking));
    HIPCHECK(hipStreamCreateWithFlags(&c->gstream2, hipStreamNonBlocking));
    for (hipEvent_t* e : {&c->ev_fork, &c->ev_join, &c->ev_gfork, &c->ev_gjoin, &c->ev_gin, &c->ev_gout})
      HIPCHECK(hipEventCreateWithFlags(e, hipEventDisableTiming));
    for (hipEvent_t& e : c->prep_ev) HIPCHECK(hipEventCreate(&e));
    return BEMFMM_OK;
  });
  if (rc != BEMFMM_OK) {
    std::fprintf(stderr, "bemfmm_ctx_create: %s\n", c->err.c_str());
    bemfmm_ctx_destroy(c);  // releases the streams and events created before the failure
    return nullptr;
  }
  return c;
}

void bemfmm_ctx_destroy(bemfmm_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (hipStream_t s : {c->stream2, c->gstream, c->gstream2})
    if (s) (void)hipStreamSynchronize(s);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto e : {c->ev_fork, c->ev_join, c->ev_gfork, c->ev_gjoin, c->ev_gin, c->ev_gout, c->prep_ev[0], c->prep_ev[1]})
    if (e) (void)hipEventDestroy(e);
  drop_graphs(c);
  if (c->op_stage) (void)pinned_free(c->op_stage);
  if (c->hres) (void)pinned_free(c->hres);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  for (hipStream_t s : {c->stream2, c->gstream, c->gstream2})
    if (s) (void)hipStreamDestroy(s);
  delete c;
}

const char* bemfmm_last_error(const bemfmm_ctx* c) { return c ? c->err.c_str() : "null context"; }

int bemfmm_set_geometry(bemfmm_ctx* c, int n_nodes, const double* X, int n_faces, const int32_t* tri, const int32_t* face_dof,
                        int n_dof, const int32_t* dof_kind, const int32_t* dof_point) {
  if (!c) return BEMFMM_E_INVALID;
  return guarded(c, [&]() {
    if (n_nodes <= 0 || n_faces <= 0 || n_dof <= 0 || !X || !tri || !face_dof || !dof_kind || !dof_point)
      throw std::runtime_error("bemfmm_set_geometry: empty or null input");
    c->n_nodes = n_nodes;
    c->n_faces = n_faces;
    c->n_dof = n_dof;
    c->X.assign(X, X + 3 * (size_t)n_nodes);
    // element_order 2: six nodes and six DOF entries per face (vertices, then edge midpoints)
    const size_t npf = c->opts.element_order == 2 ? 6 : 3;
    c->tri.assign(tri, tri + npf * (size_t)n_faces);
    c->face_dof.assign(face_dof, face_dof + npf * (size_t)n_faces);
    c->dof_kind.assign(dof_kind, dof_kind + n_dof);
    c->dof_point.assign(dof_point, dof_point + n_dof);
    c->have_geom = true;
    c->have_bc = false;
    c->prepared = false;
    c->robin_bc.clear();
    return BEMFMM_OK;
  });
}

int bemfmm_set_bc(bemfmm_ctx* c, const double* phi, const double* phin) {
  if (!c) return BEMFMM_E_INVALID;
  return guarded(c, [&]() {
    if (!c->have_geom) throw HipError("bemfmm_set_bc: geometry not set", BEMFMM_E_STATE);
    if (!phi || !phin) throw std::runtime_error("bemfmm_set_bc: null input");
    for (int i = 0; i < c->n_dof; ++i) {
      const bool dir = c->dof_kind[i] == 0;
      if ((dir || c->dof_kind[i] == 2) && !std::isfinite(phi[i])) throw HipError("non-finite phi", BEMFMM_E_NONFINITE);
      if (!dir && !std::isfinite(phin[i])) throw HipError("non-finite phin", BEMFMM_E_NONFINITE);
    }
    c->phi_bc.assign(phi, phi + c->n_dof);
    c->phin_bc.assign(phin, phin + c->n_dof);
    c->have_bc = true;
    if (c->prepared) {
      // boundary values only change the right-hand side: refresh the knowns in place
      upload_knowns(c);
    }
    return BEMFMM_OK;
  });
}

int bemfmm_set_robin(bemfmm_ctx* c, const double* kappa) {
  if (!c) return BEMFMM_E_INVALID;
  return guarded(c, [&]() {
    if (!c->have_geom) throw HipError("bemfmm_set_robin: geometry not set", BEMFMM_E_STATE);
    if (!kappa) throw std::runtime_error("bemfmm_set_robin: null input");
    for (int i = 0; i < c->n_dof; ++i)
      if (c->dof_kind[i] == 3 && !std::isfinite(kappa[i])) throw HipError("non-finite kappa", BEMFMM_E_NONFINITE);
    c->robin_bc.assign(kappa, kappa + c->n_dof);
    c->prepared = false;  // the operator and its preconditioner depend on kappa: prepare again
    return BEMFMM_OK;
  });
}

int bemfmm_prepare(bemfmm_ctx* c) {
  if (!c) return BEMFMM_E_INVALID;
  return guarded(c, [&]() {
    if (!c->have_geom || !c->have_bc) throw HipError("bemfmm_prepare: set geometry and boundary values first", BEMFMM_E_STATE);
    do_prepare(c);
    return BEMFMM_OK;
  });
}

int bemfmm_matvec(bemfmm_ctx* c, const double* x, double* y) {
  if (!c) return BEMFMM_E_INVALID;
  return guarded(c, [&]() {
    if (!c->prepared) throw HipError("bemfmm_matvec: call bemfmm_prepare first", BEMFMM_E_STATE);
    host_matvec(c, x, y, false);
    return BEMFMM_OK;
  });
}

int bemfmm_matvec_device(bemfmm_ctx* c, const double* x, double* y) {
  if (!c) return BEMFMM_E_INVALID;
  return guarded(c, [&]() {
    if (!c->prepared) throw HipError("bemfmm_matvec_device: call bemfmm_prepare first", BEMFMM_E_STATE);
    device_matvec(c, x, y, false);
    return BEMFMM_OK;
  });
}

int bemfmm_matvec_device_noalloc(bemfmm_ctx* c, const double* x, double* y) {
  if (!c) return BEMFMM_E_INVALID;
  return guarded(c, [&]() {
    if (!c->prepared) throw HipError("bemfmm_matvec_device_noalloc: call bemfmm_prepare first", BEMFMM_E_STATE);
    if (!x || !y) throw std::runtime_error("bemfmm_matvec_device_noalloc: null vector");
    void* base = nullptr;
    size_t cap = 0;
    HIPCHECK(dev::scratch_region(&base, &cap));
    // work buffers carved from the shared per-GPU array, each 256-byte aligned
    const int64_t n = c->n_dof, nbC = (int64_t)c->n_boxes * c->C;
    size_t off = 0;
    auto carve = [&](int64_t count) {
      double* p = reinterpret_cast<double*>(static_cast<char*>(base) + off);
      off += ((size_t)count * sizeof(double) + 255) & ~size_t(255);
      return p;
    };
    MatvecWork w = context_work(c);  // the context's streams and events; its buffers are replaced below
    w.stage = nullptr;
    w.m32 = nullptr;
    w.buf0 = carve(n + nbC);
    w.buf1 = carve(n + nbC);
    w.M0 = carve(nbC);
    w.M1 = carve(nbC);
    w.y0 = carve(c->n_nodes);
    w.y1 = carve(c->n_nodes);
    double* xi = c->perm_identity ? nullptr : carve(n);
    double* yi = c->perm_identity ? nullptr : carve(n);
    if (off > cap)
      throw HipError("bemfmm_matvec_device_noalloc: this problem needs " + std::to_string((off + (1 << 20) - 1) >> 20) +
                         " MiB of work memory, more than the " + std::to_string(cap >> 20) +
                         " MiB reserved per GPU for allocation-free calls (use bemfmm_matvec_device)",
                     BEMFMM_E_INVALID);
    c->n_matvec++;
    if (c->perm_identity) {
      evaluate_launch(c, 0, x, y, false, w);
    } else {
      dev::perm_gather(c->stream, (int)n, c->d_perm.p, x, xi);
      evaluate_launch(c, 0, xi, yi, false, w);
      dev::perm_scatter(c->stream, (int)n, c->d_perm.p, yi, y);
    }
    HIPCHECK(hipGetLastError());
    return BEMFMM_OK;
  });
}

int bemfmm_matvec_fp32(bemfmm_ctx* c, const double* x, double* y) {
  if (!c) return BEMFMM_E_INVALID;
  return guarded(c, [&]() {
    if (!c->prepared) throw HipError("bemfmm_matvec_fp32: call bemfmm_prepare first", BEMFMM_E_STATE);
    if (c->opts.precision != BEMFMM_PRECISION_FP32_REFINE)
      throw HipError("bemfmm_matvec_fp32: context not created with BEMFMM_PRECISION_FP32_REFINE", BEMFMM_E_STATE);
    host_matvec(c, x, y, true);
    return BEMFMM_OK;
  });
}

int bemfmm_matvec_fp32_device(bemfmm_ctx* c, const double* x, double* y) {
  if (!c) return BEMFMM_E_INVALID;
  return guarded(c, [&]() {
    if (!c->prepared) throw HipError("bemfmm_matvec_fp32_device: call bemfmm_prepare first", BEMFMM_E_STATE);
    if (c->opts.precision != BEMFMM_PRECISION_FP32_REFINE)
      throw HipError("bemfmm_matvec_fp32_device: context not created with BEMFMM_PRECISION_FP32_REFINE", BEMFMM_E_STATE);
    device_matvec(c, x, y, true);
    return BEMFMM_OK;
  });
}

int bemfmm_sync(bemfmm_ctx* c) {
  if (!c) return BEMFMM_E_INVALID;
  return guarded(c, [&]() {
    HIPCHECK(hipStreamSynchronize(c->stream));
    return BEMFMM_OK;
  });
}

int bemfmm_rhs(bemfmm_ctx* c, double* b) {
  if (!c) return BEMFMM_E_INVALID;
  return guarded(c, [&]() {
    if (!c->prepared) throw HipError("bemfmm_rhs: call bemfmm_prepare first", BEMFMM_E_STATE);
    evaluate(c, 1, nullptr, c->yout.p);
    download_caller(c, c->yout.p, b);
    return BEMFMM_OK;
  });
}

int bemfmm_alpha(bemfmm_ctx* c, double* a) {
  if (!c) return BEMFMM_E_INVALID;
  return guarded(c, [&]() {
    if (!c->prepared) throw HipError("bemfmm_alpha: call bemfmm_prepare first", BEMFMM_E_STATE);
    download_caller(c, c->alpha.p, a);
    return BEMFMM_OK;
  });
}

int bemfmm_solve(bemfmm_ctx* c, const double* x0, double* x, double tol, int max_iter, int restart, bemfmm_result* r) {
  if (!c) return BEMFMM_E_INVALID;
  return guarded(c, [&]() {
    if (!x0 || !x || !(tol > 0) || max_iter < 1) throw std::runtime_error("bemfmm_solve: invalid arguments");
    return do_solve(c, x0, x, tol, max_iter, restart, r);
  });
}

int bemfmm_residual_history(const bemfmm_ctx* c, double* out, int cap) {
  if (!c) return BEMFMM_E_INVALID;
  const int n = (int)c->history.size();
  if (out)
    for (int i = 0; i < std::min(n, cap); ++i) out[i] = c->history[i];
  return n;
}

int bemfmm_precond_apply(bemfmm_ctx* c, const double* r, double* z) {
  if (!c) return BEMFMM_E_INVALID;
  return guarded(c, [&]() {
    if (!c->prepared) throw HipError("bemfmm_precond_apply: call bemfmm_prepare first", BEMFMM_E_STATE);
    std::vector<double> ri(c->n_dof);
    host_to_internal(c, r, ri.data());
    HIPCHECK(hipMemcpyAsync(c->xin.p, ri.data(), c->n_dof * sizeof(double), hipMemcpyHostToDevice, c->stream));
    precond_apply(c, c->xin.p, c->yout.p);
    download_caller(c, c->yout.p, z);
    return BEMFMM_OK;
  });
}

int bemfmm_dof_order(const bemfmm_ctx* c, int32_t* order) {
  if (!c || !order) return BEMFMM_E_INVALID;
  if (!c->prepared) return BEMFMM_E_STATE;
  std::copy(c->perm.begin(), c->perm.end(), order);
  return BEMFMM_OK;
}

int bemfmm_schwarz_blocks(const bemfmm_ctx* c, int64_t* ptr, int32_t* idx, int32_t* ncore) {
  if (!c) return BEMFMM_E_INVALID;
  if (!c->prepared) return BEMFMM_E_STATE;
  const auto& B = c->sw_host;
  const int nb = B.n_blocks();
  if (ptr) std::copy(B.ptr.begin(), B.ptr.end(), ptr);
  if (idx)
    for (size_t k = 0; k < B.idx.size(); ++k) idx[k] = c->perm[B.idx[k]];  // caller numbering
  if (ncore) std::copy(B.ncore.begin(), B.ncore.end(), ncore);
  return nb;
}

int bemfmm_tree_digest(const bemfmm_ctx* c, uint64_t* digest) {
  if (!c || !digest) return BEMFMM_E_INVALID;
  if (!c->prepared || !c->plan) return BEMFMM_E_STATE;
  const Plan& pl = *c->plan;
  uint64_t h = 0x9e3779b97f4a7c15ull;
  auto mix = [&](uint64_t v) {  // splitmix64 finaliser over a running state
    h += v + 0x9e3779b97f4a7c15ull;
    uint64_t z = h;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    h = z ^ (z >> 31);
  };
  auto add_i32 = [&](const std::vector<int32_t>& v) {
    mix(v.size());
    for (int32_t x : v) mix((uint32_t)x);
  };
  auto add_i64 = [&](const std::vector<int64_t>& v) {
    mix(v.size());
    for (int64_t x : v) mix((uint64_t)x);
  };
  add_i32(pl.level_ptr);
  add_i32(pl.box_parent);
  add_i32(pl.box_leaf);
  mix(pl.box_center.size());
  for (double x : pl.box_center) {
    uint64_t u;
    std::memcpy(&u, &x, 8);
    mix(u);
  }
  add_i32(pl.face_perm);
  add_i64(pl.box_range);
  add_i32(pl.target_box);
  add_i64(pl.near_ptr);
  add_i32(pl.near_idx);
  add_i64(pl.m2l_ptr);
  add_i32(pl.m2l_idx);
  *digest = h;
  return BEMFMM_OK;
}

int bemfmm_refinement_history(const bemfmm_ctx* c, double* out, int cap) {
  if (!c) return BEMFMM_E_INVALID;
  const int n = (int)c->refine_history.size();
  if (out)
    for (int i = 0; i < std::min(n, cap); ++i) out[i] = c->refine_history[i];
  return n;
}

int bemfmm_get_stats(const bemfmm_ctx* c, bemfmm_stats* s) {
  if (!c || !s) return BEMFMM_E_INVALID;
  if (!c->prepared) return BEMFMM_E_STATE;
  *s = c->stats;
  return BEMFMM_OK;
}

int bemfmm_matvec_profile(bemfmm_ctx* c, double* ms8) {
  if (!c) return BEMFMM_E_INVALID;
  if (!ms8) {  // turn event recording off (matvecs replay their graph again)
    c->prof = false;
    return BEMFMM_OK;
  }
  return guarded(c, [&]() {
    if (!c->prof) {
      for (auto& e : c->ev) HIPCHECK(hipEventCreate(&e));
      c->prof = true;
      for (int i = 0; i < 8; ++i) ms8[i] = 0.;
      return BEMFMM_OK;
    }
    HIPCHECK(hipEventSynchronize(c->ev[8]));
    for (int i = 0; i < 8; ++i) {
      float ms = 0.f;
      HIPCHECK(hipEventElapsedTime(&ms, c->ev[i], c->ev[i + 1]));
      ms8[i] = ms;
    }
    return BEMFMM_OK;
  });
}

}  // extern "C"
