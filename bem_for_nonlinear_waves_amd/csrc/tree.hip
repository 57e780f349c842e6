// Device kernels of the plan's tree stage: the octree (generateTree, lib/include/lib_spatial_partitioning.hpp:
// 1079-1115, with the grow condition of bem/BEM_solveBVP_GMRES_FMM.hpp:1199), the near and M2L interaction
// lists (buildNearAndM2LLists, lib/include/moments.hpp:968-997) and the deepest box of every target node
// (getBucketAtDeepest), level-synchronously over faces and box pairs, with rocPRIM radix sorts on path keys.
//
// A face's path key holds one 5-bit digit per level: 1 + its cell in the box it was split from, or 0 once
// the face rests in a leaf or was lost by the child's inside test (it then stays with the splitting box, as
// in the reference). Sorting the faces by path key and face id gives the depth-first face order with every
// box's faces contiguous, lost faces ahead of the children; sorting the (box, cell, face) keys of a level
// makes the next level's boxes in depth-first order. The per-level driver is host code (bemfmm.cpp); every
// function here enqueues one kernel or one rocPRIM call on the given stream.
//
// Compiled with -ffp-contract=off: the cell arithmetic and the acceptance test (treemath.hpp) must round as
// the host plan's do, so that both build the same tree and the same lists.
#include <hip/hip_runtime.h>

#include <cstring>  // rocPRIM's texture iterator uses memset without including it

#include <rocprim/rocprim.hpp>

#include "kernels.hpp"
#include "treemath.hpp"

namespace bemfmm {
namespace dev {

namespace {

constexpr int kTB = 256;
inline unsigned grid_of(int64_t n) { return (unsigned)((n + kTB - 1) / kTB); }

// Slot of this lane in an append-only array when pred holds: one atomic per wave. Every lane of the wave
// must call it (no lane may have returned before).
__device__ inline uint32_t wave_append(uint32_t* counter, bool pred) {
  const uint64_t mask = __ballot(pred);
  const int lane = __lane_id();
  const int leader = mask ? __ffsll((unsigned long long)mask) - 1 : 0;
  uint32_t base = 0;
  if (mask && lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
  base = __shfl(base, leader);
  return base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
}

__device__ inline int64_t lower_bound_u64(const uint64_t* a, int64_t n, uint64_t v) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ void k_tree_split(int F, int lb, int leaf_max, int may_split, const double* __restrict__ pole,
                             const int32_t* __restrict__ fbox, uint64_t* __restrict__ fkey,
                             const BoxGeom* __restrict__ geom, const int32_t* __restrict__ count,
                             uint64_t* __restrict__ out, uint32_t* counter, int32_t* err) {
  const int f = blockIdx.x * kTB + threadIdx.x;
  bool split = false;
  uint64_t key = 0;
  if (f < F) {
    fkey[f] <<= 5;
    const int b = fbox[f];
    if (b >= lb && may_split && count[b] > leaf_max) {
      const double x[3] = {pole[3 * f], pole[3 * f + 1], pole[3 * f + 2]};
      const int cell = box_cell_of(geom[b], x);
      if (cell >= kTreeCells) {
        atomicOr(err, 1);
      } else {
        split = true;
        key = ((uint64_t)((uint32_t)(b - lb) * 32u + (uint32_t)cell) << 32) | (uint32_t)f;
      }
    }
  }
  const uint32_t slot = wave_append(counter, split);
  if (split) out[slot] = key;
}

__global__ void k_tree_heads(int n, const uint64_t* __restrict__ keys, int32_t* __restrict__ flags) {
  const int i = blockIdx.x * kTB + threadIdx.x;
  if (i >= n) return;
  flags[i] = (i == 0 || (keys[i] >> 32) != (keys[i - 1] >> 32)) ? 1 : 0;
}

__global__ void k_tree_children(int n, const uint64_t* __restrict__ keys, const int32_t* __restrict__ ci, int lb,
                                int le, BoxGeom* geom, int32_t* parent, int32_t* level, uint64_t* pk, int32_t* count,
                                int32_t* nchild, int32_t* first_child, int32_t* child_tab) {
  const int i = blockIdx.x * kTB + threadIdx.x;
  if (i >= n) return;
  const uint32_t hi = (uint32_t)(keys[i] >> 32);
  if (i > 0 && (uint32_t)(keys[i - 1] >> 32) == hi) return;  // not the first face of its cell
  const int pb = (int)(hi / 32u), cell = (int)(hi % 32u);
  const int P = lb + pb, c = le + ci[i] - 1;
  BoxGeom g;
  box_child(geom[P], cell, g);
  geom[c] = g;
  parent[c] = P;
  level[c] = level[P] + 1;
  pk[c] = pk[P] * 32u + (uint64_t)(cell + 1);
  count[c] = 0;
  nchild[c] = 0;
  first_child[c] = -1;
  for (int k = 0; k < kTreeCells; ++k) child_tab[(int64_t)c * kTreeCells + k] = -1;
  child_tab[(int64_t)P * kTreeCells + cell] = c;
  atomicAdd(&nchild[P], 1);
  if (i == 0 || (uint32_t)(keys[i - 1] >> 32) / 32u != (uint32_t)pb) first_child[P] = c;
}

__global__ void k_tree_assign(int n, const uint64_t* __restrict__ keys, const int32_t* __restrict__ ci, int le,
                              const double* __restrict__ pole, const BoxGeom* __restrict__ geom,
                              int32_t* __restrict__ fbox, uint64_t* __restrict__ fkey, int32_t* count) {
  const int i = blockIdx.x * kTB + threadIdx.x;
  int c = -1;
  bool ins = false;
  if (i < n) {
    const int f = (int)(uint32_t)keys[i];
    const int cell = (int)((uint32_t)(keys[i] >> 32) % 32u);
    c = le + ci[i] - 1;
    const double x[3] = {pole[3 * f], pole[3 * f + 1], pole[3 * f + 2]};
    ins = box_inside(geom[c], x);
    if (ins) {
      fbox[f] = c;
      fkey[f] |= (uint64_t)(cell + 1);
    }
  }
  // the faces of one child are consecutive (sorted keys): the last lane of each run in the wave adds the run's
  // count, one atomic per run instead of one per face
  const int lane = __lane_id();
  const uint64_t insm = __ballot(ins);
  const int cprev = __shfl_up(c, 1), cnext = __shfl_down(c, 1);
  const bool head = lane == 0 || cprev != c;
  const bool tail = lane == 63 || cnext != c;
  const uint64_t heads = __ballot(head);
  if (tail && c >= 0) {
    const uint64_t upto = (2ull << lane) - 1ull;  // lanes 0 .. lane (all ones for lane 63)
    const int start = 63 - __clzll(heads & upto);
    const int cnt = __popcll(insm & upto & ~((1ull << start) - 1ull));
    if (cnt) atomicAdd(&count[c], cnt);
  }
}

__global__ void k_tree_nonempty(int nc, int le, const int32_t* __restrict__ count, int32_t* err) {
  const int i = blockIdx.x * kTB + threadIdx.x;
  if (i < nc && count[le + i] == 0) atomicOr(err, 2);
}

template <typename T>
__global__ void k_fill_int(int64_t n, T* p, T v) {
  const int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_iota(int n, int32_t* out) {
  const int i = blockIdx.x * kTB + threadIdx.x;
  if (i < n) out[i] = i;
}

__global__ void k_tree_ranges(int nb, int digits, const int32_t* __restrict__ level, const uint64_t* __restrict__ pk,
                              const uint64_t* __restrict__ keys, int F, int64_t* __restrict__ range) {
  const int b = blockIdx.x * kTB + threadIdx.x;
  if (b >= nb) return;
  const int sh = 5 * (digits - level[b]);
  range[2 * b] = lower_bound_u64(keys, F, pk[b] << sh);
  range[2 * b + 1] = lower_bound_u64(keys, F, (pk[b] + 1) << sh);
}

__global__ void k_tree_leaf_off(int nb, const int32_t* __restrict__ nchild, const int64_t* __restrict__ range,
                                const int32_t* __restrict__ perm, const double* __restrict__ maxoff,
                                double* __restrict__ off) {
  const int b = blockIdx.x * kTB + threadIdx.x;
  if (b >= nb) return;
  double o = 0.;
  if (nchild[b] == 0)
    for (int64_t i = range[2 * b]; i < range[2 * b + 1]; ++i) {
      const double v = maxoff[perm[i]];
      o = o < v ? v : o;
    }
  off[b] = o;
}

__global__ void k_tree_parent_off(int lb, int le, const int32_t* __restrict__ nchild,
                                  const int32_t* __restrict__ first_child, double* off) {
  const int b = lb + blockIdx.x * kTB + threadIdx.x;
  if (b >= le || nchild[b] == 0) return;
  double o = 0.;
  for (int c = first_child[b]; c < first_child[b] + nchild[b]; ++c) o = o < off[c] ? off[c] : o;
  off[b] = o;
}

__global__ void k_tree_targets(int N, const double* __restrict__ X, const BoxGeom* __restrict__ geom,
                               const int32_t* __restrict__ nchild, const int32_t* __restrict__ child_tab,
                               int32_t* __restrict__ target) {
  const int i = blockIdx.x * kTB + threadIdx.x;
  if (i >= N) return;
  const double x[3] = {X[3 * i], X[3 * i + 1], X[3 * i + 2]};
  int cur = 0;
  while (nchild[cur] > 0) {
    const int cell = box_cell_of(geom[cur], x);
    const int nx = cell < kTreeCells ? child_tab[(int64_t)cur * kTreeCells + cell] : -1;
    if (nx < 0) break;
    cur = nx;
  }
  target[i] = cur;
}

// per near-candidate pair (A, B): near flag, and the counts of its child pairs that are M2L pairs and that
// are next level's near candidates
__global__ void k_list_count(int np, const uint64_t* __restrict__ pairs, const BoxGeom* __restrict__ geom,
                             const double* __restrict__ off, const int32_t* __restrict__ nchild,
                             const int32_t* __restrict__ first_child, double theta, int32_t* __restrict__ nearf,
                             int32_t* __restrict__ m2lc, int32_t* __restrict__ nextc) {
  const int i = blockIdx.x * kTB + threadIdx.x;
  if (i >= np) return;
  const int A = (int)(pairs[i] >> 32), B = (int)(uint32_t)pairs[i];
  int nf = 0, nm = 0, nn = 0;
  if (!boxes_far(geom[A].X, geom[A].dL, off[A], geom[B].X, geom[B].dL, off[B], theta)) {
    nf = A != B;
    const int a0 = first_child[A], a1 = a0 + nchild[A], b0 = first_child[B], b1 = b0 + nchild[B];
    for (int ac = a0; ac < a1; ++ac)
      for (int bc = b0; bc < b1; ++bc) {
        if (ac != bc && boxes_far(geom[ac].X, geom[ac].dL, off[ac], geom[bc].X, geom[bc].dL, off[bc], theta)) ++nm;
        else ++nn;
      }
  }
  nearf[i] = nf;
  m2lc[i] = nm;
  nextc[i] = nn;
}

__global__ void k_list_write(int np, const uint64_t* __restrict__ pairs, const BoxGeom* __restrict__ geom,
                             const double* __restrict__ off, const int32_t* __restrict__ nchild,
                             const int32_t* __restrict__ first_child, double theta,
                             const int32_t* __restrict__ nearf, const int32_t* __restrict__ near_at,
                             const int32_t* __restrict__ m2l_at, const int32_t* __restrict__ next_at,
                             uint64_t* __restrict__ near_out, uint64_t* __restrict__ m2l_out,
                             uint64_t* __restrict__ next_out) {
  const int i = blockIdx.x * kTB + threadIdx.x;
  if (i >= np) return;
  const int A = (int)(pairs[i] >> 32), B = (int)(uint32_t)pairs[i];
  if (nearf[i]) near_out[near_at[i]] = pairs[i];
  if (boxes_far(geom[A].X, geom[A].dL, off[A], geom[B].X, geom[B].dL, off[B], theta)) return;
  int64_t km = m2l_at[i], kn = next_at[i];
  const int a0 = first_child[A], a1 = a0 + nchild[A], b0 = first_child[B], b1 = b0 + nchild[B];
  for (int ac = a0; ac < a1; ++ac)
    for (int bc = b0; bc < b1; ++bc) {
      const uint64_t key = ((uint64_t)(uint32_t)ac << 32) | (uint32_t)bc;
      if (ac != bc && boxes_far(geom[ac].X, geom[ac].dL, off[ac], geom[bc].X, geom[bc].dL, off[bc], theta))
        m2l_out[km++] = key;
      else
        next_out[kn++] = key;
    }
}

// totals of the three exclusive scans: out[k] = at_k[np-1] + count_k[np-1]
__global__ void k_list_totals(int np, const int32_t* nearf, const int32_t* m2lc, const int32_t* nextc,
                              const int32_t* near_at, const int32_t* m2l_at, const int32_t* next_at, int64_t* out) {
  if (threadIdx.x != 0) return;
  out[0] = (int64_t)near_at[np - 1] + nearf[np - 1];
  out[1] = (int64_t)m2l_at[np - 1] + m2lc[np - 1];
  out[2] = (int64_t)next_at[np - 1] + nextc[np - 1];
}

__global__ void k_csr_from_keys(int nb, const uint64_t* __restrict__ keys, int64_t n, int64_t* __restrict__ ptr,
                                int32_t* __restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (i <= nb) ptr[i] = lower_bound_u64(keys, n, (uint64_t)i << 32);
  if (i < n) idx[i] = (int32_t)(uint32_t)keys[i];
}

__global__ void k_node_keys(int N, const int32_t* __restrict__ target, uint64_t* __restrict__ keys) {
  const int i = blockIdx.x * kTB + threadIdx.x;
  if (i < N) keys[i] = ((uint64_t)(uint32_t)target[i] << 32) | (uint32_t)i;
}

// boxes holding target nodes, in box order: flags for the scan, then the compacted list
__global__ void k_ns_tbox_flags(int nb, const int64_t* __restrict__ tptr, int32_t* __restrict__ flags) {
  const int b = blockIdx.x * kTB + threadIdx.x;
  if (b < nb) flags[b] = tptr[b + 1] > tptr[b] ? 1 : 0;
}
__global__ void k_ns_tbox_compact(int nb, const int32_t* __restrict__ flags, const int32_t* __restrict__ at,
                                  int32_t* __restrict__ tbox) {
  const int b = blockIdx.x * kTB + threadIdx.x;
  if (b < nb && flags[b]) tbox[at[b] - 1] = b;
}
// chunks of <= kNearRows rows per target box (entry ntb: 0, so that the exclusive scan ends with the total)
__global__ void k_ns_chunk_count(int ntb, const int64_t* __restrict__ tptr, const int32_t* __restrict__ tbox,
                                 int32_t* __restrict__ cnt) {
  const int ti = blockIdx.x * kTB + threadIdx.x;
  if (ti > ntb) return;
  cnt[ti] = ti == ntb ? 0 : (int32_t)((tptr[tbox[ti] + 1] - tptr[tbox[ti]] + kNearRows - 1) / kNearRows);
}

// ---- near-field blocks (plan.cpp plan_finish, "near blocks") ----
// ranges per target box ti (box b): b itself, b's near boxes, and the near leaves of b's ancestors
__global__ void k_ns_range_count(int ntb, const int32_t* __restrict__ tbox, const int32_t* __restrict__ parent,
                                 const int64_t* __restrict__ near_ptr, const int32_t* __restrict__ near_idx,
                                 const int32_t* __restrict__ nchild, int32_t* __restrict__ cnt) {
  const int ti = blockIdx.x * kTB + threadIdx.x;
  if (ti > ntb) return;
  if (ti == ntb) {  // 0, so that the exclusive scan ends with the total
    cnt[ti] = 0;
    return;
  }
  const int b = tbox[ti];
  int n = 1 + (int)(near_ptr[b + 1] - near_ptr[b]);
  for (int a = parent[b]; a >= 0; a = parent[a])
    for (int64_t k = near_ptr[a]; k < near_ptr[a + 1]; ++k) n += nchild[near_idx[k]] == 0;
  cnt[ti] = n;
}

__global__ void k_ns_range_write(int ntb, const int32_t* __restrict__ tbox, const int32_t* __restrict__ parent,
                                 const int64_t* __restrict__ near_ptr, const int32_t* __restrict__ near_idx,
                                 const int32_t* __restrict__ nchild, const int64_t* __restrict__ box_range,
                                 const int32_t* __restrict__ at, uint64_t* __restrict__ keys,
                                 int32_t* __restrict__ ends) {
  const int ti = blockIdx.x * kTB + threadIdx.x;
  if (ti >= ntb) return;
  const int b = tbox[ti];
  int k = at[ti];
  auto put = [&](int box) {
    keys[k] = ((uint64_t)(uint32_t)ti << 32) | (uint32_t)box_range[2 * box];
    ends[k] = (int32_t)box_range[2 * box + 1];
    ++k;
  };
  put(b);
  for (int64_t q = near_ptr[b]; q < near_ptr[b + 1]; ++q) put(near_idx[q]);
  for (int a = parent[b]; a >= 0; a = parent[a])
    for (int64_t q = near_ptr[a]; q < near_ptr[a + 1]; ++q)
      if (nchild[near_idx[q]] == 0) put(near_idx[q]);
}

// merges each target box's sorted ranges in place (empty ones dropped, touching ones joined); err |= 4 on an
// overlap. Writes the merged count and the face count of the target box.
__global__ void k_ns_range_merge(int ntb, const int32_t* __restrict__ at, const int32_t* __restrict__ cnt,
                                 const uint64_t* __restrict__ keys, const int32_t* __restrict__ ends,
                                 int64_t* __restrict__ mstart, int64_t* __restrict__ mend, int32_t* __restrict__ mcnt,
                                 int64_t* __restrict__ nfaces, int32_t* err) {
  const int ti = blockIdx.x * kTB + threadIdx.x;
  if (ti > ntb) return;
  if (ti == ntb) {
    mcnt[ti] = 0;
    nfaces[ti] = 0;
    return;
  }
  const int k0 = at[ti], k1 = k0 + cnt[ti];
  int m = 0;
  int64_t nf = 0;
  bool overlap = false;
  for (int k = k0; k < k1; ++k) {
    const int64_t a = (int64_t)(uint32_t)keys[k], e = ends[k];
    if (a == e) continue;
    if (m > 0 && a <= mend[k0 + m - 1]) {
      if (a < mend[k0 + m - 1]) overlap = true;
      if (e > mend[k0 + m - 1]) mend[k0 + m - 1] = e;
    } else {
      mstart[k0 + m] = a;
      mend[k0 + m] = e;
      ++m;
    }
  }
  for (int q = 0; q < m; ++q) nf += mend[k0 + q] - mstart[k0 + q];
  if (overlap) atomicOr(err, 4);
  mcnt[ti] = m;
  nfaces[ti] = nf;
}

// one key (target box, DOF) per face node of each target box's merged ranges; fdof in face_perm order, npf per
// face (an edge midpoint without a DOF contributes its first end vertex's DOF, which the face has anyway)
__global__ void k_ns_dof_keys(int npf, int64_t total, int ntb, const int64_t* __restrict__ fstart,
                              const int32_t* __restrict__ at, const int32_t* __restrict__ mcnt,
                              const int64_t* __restrict__ mstart, const int64_t* __restrict__ mend,
                              const int32_t* __restrict__ fdof, uint64_t* __restrict__ keys) {
  const int64_t e = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (e >= total) return;
  int lo = 0, hi = ntb - 1;  // last target box with fstart <= e
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (fstart[mid] <= e) lo = mid;
    else hi = mid - 1;
  }
  const int ti = lo;
  int64_t i = e - fstart[ti];
  int64_t pos = -1;
  for (int q = at[ti]; q < at[ti] + mcnt[ti]; ++q) {
    const int64_t len = mend[q] - mstart[q];
    if (i < len) {
      pos = mstart[q] + i;
      break;
    }
    i -= len;
  }
  for (int c = 0; c < npf; ++c) {
    int d = fdof[npf * pos + c];
    if (d < 0) d = fdof[npf * pos + (c - 3)];
    keys[npf * e + c] = ((uint64_t)(uint32_t)ti << 32) | (uint32_t)d;
  }
}

// the local DOF column of every face corner of every target box's merged ranges (in the near quadrature's local
// face order), so that the quadrature of each row reads three columns per face instead of searching them
__global__ void k_ns_face_cols(int npf, int64_t total, int ntb, const int64_t* __restrict__ fstart,
                               const int32_t* __restrict__ at, const int32_t* __restrict__ mcnt,
                               const int64_t* __restrict__ mstart, const int64_t* __restrict__ mend,
                               const int32_t* __restrict__ fdof, const int64_t* __restrict__ dptr,
                               const int32_t* __restrict__ dofs, int32_t* __restrict__ cols) {
  const int64_t e = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (e >= total) return;
  int lo = 0, hi = ntb - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (fstart[mid] <= e) lo = mid;
    else hi = mid - 1;
  }
  const int ti = lo;
  int64_t i = e - fstart[ti];
  int64_t pos = -1;
  for (int q = at[ti]; q < at[ti] + mcnt[ti]; ++q) {
    const int64_t len = mend[q] - mstart[q];
    if (i < len) {
      pos = mstart[q] + i;
      break;
    }
    i -= len;
  }
  const int32_t* d = dofs + dptr[ti];
  const int n = (int)(dptr[ti + 1] - dptr[ti]);
  for (int c = 0; c < npf; ++c) {
    const int v = fdof[npf * pos + c];
    if (v < 0) {  // edge midpoint without a DOF
      cols[npf * e + c] = -1;
      continue;
    }
    int a = 0, b = n - 1;
    while (a < b) {
      const int mid = (a + b) >> 1;
      if (d[mid] < v) a = mid + 1;
      else b = mid;
    }
    cols[npf * e + c] = a;
  }
}

__global__ void k_unique_flags(int64_t n, const uint64_t* __restrict__ keys, int32_t* __restrict__ flags) {
  const int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (i < n) flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

__global__ void k_compact_u64(int64_t n, const uint64_t* __restrict__ keys, const int32_t* __restrict__ flags,
                              const int32_t* __restrict__ at, uint64_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (i < n && flags[i]) out[at[i] - 1] = keys[i];
}

// chunks of <= kNearRows rows per target box, keyed for a stable sort by descending cost (rows x padded columns)
__global__ void k_ns_chunk_keys(int ntb, const int64_t* __restrict__ tptr, const int32_t* __restrict__ tbox,
                                const int32_t* __restrict__ chunk_at, const int64_t* __restrict__ dptr, int C,
                                uint64_t* __restrict__ keys, int32_t* err) {
  const int ti = blockIdx.x * kTB + threadIdx.x;
  if (ti >= ntb) return;
  const int64_t nt = tptr[tbox[ti] + 1] - tptr[tbox[ti]];
  const int64_t ncol = ((dptr[ti + 1] - dptr[ti]) + C + 3) & ~(int64_t)3;
  int k = chunk_at[ti];
  for (int64_t r0 = 0; r0 < nt; r0 += kNearRows, ++k) {
    const int64_t rows = (nt - r0 < kNearRows) ? nt - r0 : kNearRows;
    const int64_t cost = rows * ncol;
    if (cost >= (1 << 22)) atomicOr(err, 8);
    keys[k] = ((uint64_t)((1 << 22) - 1 - cost) << 24) | (uint32_t)k;
  }
}

// per chunk in cost order: its target box, first row, rows, columns, DOF columns, ranges
__global__ void k_ns_chunk_sizes(int nch, const uint64_t* __restrict__ order, const int32_t* __restrict__ chunk_ti,
                                 const int32_t* __restrict__ chunk_at, const int64_t* __restrict__ tptr,
                                 const int32_t* __restrict__ tbox, const int64_t* __restrict__ dptr,
                                 const int32_t* __restrict__ mcnt, int C, int64_t* __restrict__ rows_v,
                                 int64_t* __restrict__ cols_v, int64_t* __restrict__ mat_v, int64_t* __restrict__ rng_v,
                                 int64_t* __restrict__ nnz_v) {
  const int k = blockIdx.x * kTB + threadIdx.x;
  if (k > nch) return;
  if (k == nch) {  // the exclusive scans' totals land at index nch
    rows_v[k] = cols_v[k] = mat_v[k] = rng_v[k] = nnz_v[k] = 0;
    return;
  }
  const int c = (int)(order[k] & 0xFFFFFFu);  // the chunk index: the key's low 24 bits
  const int ti = chunk_ti[c];
  const int64_t r0 = (int64_t)(c - chunk_at[ti]) * kNearRows;
  const int64_t nt = tptr[tbox[ti] + 1] - tptr[tbox[ti]];
  const int64_t rows = (nt - r0 < kNearRows) ? nt - r0 : kNearRows;
  const int64_t ncd = dptr[ti + 1] - dptr[ti];
  const int64_t ncol = (ncd + C + 3) & ~(int64_t)3;
  rows_v[k] = rows;
  cols_v[k] = ncol;
  mat_v[k] = rows * 2 * ncol;
  rng_v[k] = mcnt[ti];
  nnz_v[k] = rows * ncd;
}

__global__ void k_ns_chunk_ti(int ntb, const int32_t* __restrict__ chunk_at, int32_t* __restrict__ chunk_ti) {
  const int ti = blockIdx.x * kTB + threadIdx.x;
  if (ti >= ntb) return;
  for (int k = chunk_at[ti]; k < chunk_at[ti + 1]; ++k) chunk_ti[k] = ti;
}

// one workgroup per chunk: its rows (target nodes in node order), columns (the target box's DOFs, its C L2P
// columns, zero-weight padding), ranges with their face prefixes, and the row -> chunk / node -> row maps
__global__ void k_ns_fill(int nch, const uint64_t* __restrict__ order, const int32_t* __restrict__ chunk_ti,
                          const int32_t* __restrict__ chunk_at, const int64_t* __restrict__ tptr,
                          const int32_t* __restrict__ tnodes, const int32_t* __restrict__ tbox,
                          const int64_t* __restrict__ dptr, const int32_t* __restrict__ dofs,
                          const int32_t* __restrict__ rat, const int32_t* __restrict__ mcnt,
                          const int64_t* __restrict__ mstart, const int64_t* __restrict__ mend, int C, int n_dof,
                          const int64_t* __restrict__ row_ptr, const int64_t* __restrict__ col_ptr,
                          const int64_t* __restrict__ range_ptr, int32_t* __restrict__ row_out,
                          int32_t* __restrict__ row_block, int64_t* __restrict__ node_row, int32_t* __restrict__ col_idx,
                          int32_t* __restrict__ ncol_dof, int64_t* __restrict__ ranges,
                          int64_t* __restrict__ range_prefix, const int64_t* __restrict__ fstart,
                          int64_t* __restrict__ block_fstart) {
  const int k = blockIdx.x;
  const int c = (int)(order[k] & 0xFFFFFFu);  // the chunk index: the key's low 24 bits
  const int ti = chunk_ti[c];
  const int b = tbox[ti];
  const int64_t r0 = (int64_t)(c - chunk_at[ti]) * kNearRows;
  const int64_t rows = row_ptr[k + 1] - row_ptr[k];
  const int64_t ncd = dptr[ti + 1] - dptr[ti];
  const int64_t ncol = col_ptr[k + 1] - col_ptr[k];
  for (int64_t r = threadIdx.x; r < rows; r += blockDim.x) {
    const int node = tnodes[tptr[b] + r0 + r];
    row_out[row_ptr[k] + r] = node;
    row_block[row_ptr[k] + r] = k;
    node_row[node] = row_ptr[k] + r;
  }
  int32_t* ci = col_idx + col_ptr[k];
  for (int64_t j = threadIdx.x; j < ncol; j += blockDim.x)
    ci[j] = j < ncd ? dofs[dptr[ti] + j] : (int32_t)(n_dof + (int64_t)b * C + (j < ncd + C ? j - ncd : C - 1));
  if (threadIdx.x == 0) {
    ncol_dof[k] = (int32_t)ncd;
    block_fstart[k] = fstart[ti];
    int64_t pre = 0, q = range_ptr[k];
    for (int g = rat[ti]; g < rat[ti] + mcnt[ti]; ++g, ++q) {
      ranges[2 * q] = mstart[g];
      ranges[2 * q + 1] = mend[g];
      range_prefix[q] = pre;
      pre += mend[g] - mstart[g];
    }
  }
}

// L2P coefficients per near row: (1/s) Re sum R_n^m(x~) L~_n^m, x~ = (x - c)/s, as plan.cpp computes them (same
// recurrences, same order of operations): thread (row, m) runs the diagonal recurrence to NR_m^m, then n > m
__global__ void k_l2p_coef(int64_t nrows, int P, const int32_t* __restrict__ row_out, const int32_t* __restrict__ target,
                           const double* __restrict__ X, const BoxGeom* __restrict__ geom,
                           const int32_t* __restrict__ level, const double* __restrict__ level_side,
                           const double* __restrict__ cnm, double* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (e >= nrows * (P + 1)) return;
  const int64_t r = e / (P + 1);
  const int m = (int)(e - r * (P + 1));
  const int C = (P + 1) * (P + 1);
  const int node = row_out[r];
  const int b = target[node];
  const double s = level_side[level[b]];
  const double x = (X[3 * node] - geom[b].X[0]) / s;
  const double y = (X[3 * node + 1] - geom[b].X[1]) / s;
  const double z = (X[3 * node + 2] - geom[b].X[2]) / s;
  const double r2 = x * x + y * y + z * z;
  double pr = 1.0, pi = 0.0;
  for (int mm = 1; mm <= m; ++mm) {
    const double sc = 1.0 / (2.0 * mm);
    const double nr = sc * (x * pr - y * pi), ni = sc * (x * pi + y * pr);
    pr = nr;
    pi = ni;
  }
  double* o = out + r * C;
  auto put = [&](int n, double re, double im) {
    const double c = cnm[n * (n + 1) / 2 + m];
    const double Rr = c * re, Ri = c * im;
    if (m == 0) {
      o[n * n] = Rr / s;
    } else {
      o[n * n + 2 * m - 1] = 2.0 * Rr / s;
      o[n * n + 2 * m] = -2.0 * Ri / s;
    }
  };
  put(m, pr, pi);
  double r1 = pr, i1 = pi, r0 = 0., i0 = 0.;  // NR_{n-1}^m, NR_{n-2}^m
  for (int n = m + 1; n <= P; ++n) {
    double ar = (2 * n - 1) * z * r1, ai = (2 * n - 1) * z * i1;
    if (n - 2 >= m) {
      ar -= r2 * r0;
      ai -= r2 * i0;
    }
    const double sc = 1.0 / ((double)(n + m) * (n - m));
    const double nr = sc * ar, ni = sc * ai;
    r0 = r1;
    i0 = i1;
    r1 = nr;
    i1 = ni;
    put(n, nr, ni);
  }
}

}  // namespace

void tree_split(hipStream_t s, int F, int lb, int leaf_max, int may_split, const double* pole, const int32_t* fbox,
                uint64_t* fkey, const BoxGeom* geom, const int32_t* count, uint64_t* out, uint32_t* counter,
                int32_t* err) {
  hipLaunchKernelGGL(k_tree_split, dim3(grid_of(F)), dim3(kTB), 0, s, F, lb, leaf_max, may_split, pole, fbox, fkey,
                     geom, count, out, counter, err);
}

void tree_heads(hipStream_t s, int n, const uint64_t* keys, int32_t* flags) {
  hipLaunchKernelGGL(k_tree_heads, dim3(grid_of(n)), dim3(kTB), 0, s, n, keys, flags);
}

void tree_children(hipStream_t s, int n, const uint64_t* keys, const int32_t* ci, int lb, int le, BoxGeom* geom,
                   int32_t* parent, int32_t* level, uint64_t* pk, int32_t* count, int32_t* nchild,
                   int32_t* first_child, int32_t* child_tab) {
  hipLaunchKernelGGL(k_tree_children, dim3(grid_of(n)), dim3(kTB), 0, s, n, keys, ci, lb, le, geom, parent, level,
                     pk, count, nchild, first_child, child_tab);
}

void tree_assign(hipStream_t s, int n, const uint64_t* keys, const int32_t* ci, int le, const double* pole,
                 const BoxGeom* geom, int32_t* fbox, uint64_t* fkey, int32_t* count) {
  hipLaunchKernelGGL(k_tree_assign, dim3(grid_of(n)), dim3(kTB), 0, s, n, keys, ci, le, pole, geom, fbox, fkey, count);
}

void tree_nonempty(hipStream_t s, int nc, int le, const int32_t* count, int32_t* err) {
  hipLaunchKernelGGL(k_tree_nonempty, dim3(grid_of(nc)), dim3(kTB), 0, s, nc, le, count, err);
}

void fill_i32(hipStream_t s, int64_t n, int32_t* p, int32_t v) {
  if (n > 0) hipLaunchKernelGGL(k_fill_int<int32_t>, dim3(grid_of(n)), dim3(kTB), 0, s, n, p, v);
}

void fill_u64(hipStream_t s, int64_t n, uint64_t* p, uint64_t v) {
  if (n > 0) hipLaunchKernelGGL(k_fill_int<uint64_t>, dim3(grid_of(n)), dim3(kTB), 0, s, n, p, v);
}

void iota_i32(hipStream_t s, int n, int32_t* out) {
  hipLaunchKernelGGL(k_iota, dim3(grid_of(n)), dim3(kTB), 0, s, n, out);
}

void tree_ranges(hipStream_t s, int nb, int digits, const int32_t* level, const uint64_t* pk, const uint64_t* keys,
                 int F, int64_t* range) {
  hipLaunchKernelGGL(k_tree_ranges, dim3(grid_of(nb)), dim3(kTB), 0, s, nb, digits, level, pk, keys, F, range);
}

void tree_leaf_off(hipStream_t s, int nb, const int32_t* nchild, const int64_t* range, const int32_t* perm,
                   const double* maxoff, double* off) {
  hipLaunchKernelGGL(k_tree_leaf_off, dim3(grid_of(nb)), dim3(kTB), 0, s, nb, nchild, range, perm, maxoff, off);
}

void tree_parent_off(hipStream_t s, int lb, int le, const int32_t* nchild, const int32_t* first_child, double* off) {
  if (le <= lb) return;
  hipLaunchKernelGGL(k_tree_parent_off, dim3(grid_of(le - lb)), dim3(kTB), 0, s, lb, le, nchild, first_child, off);
}

void tree_targets(hipStream_t s, int N, const double* X, const BoxGeom* geom, const int32_t* nchild,
                  const int32_t* child_tab, int32_t* target) {
  hipLaunchKernelGGL(k_tree_targets, dim3(grid_of(N)), dim3(kTB), 0, s, N, X, geom, nchild, child_tab, target);
}

void list_count(hipStream_t s, int np, const uint64_t* pairs, const BoxGeom* geom, const double* off,
                const int32_t* nchild, const int32_t* first_child, double theta, int32_t* nearf, int32_t* m2lc,
                int32_t* nextc) {
  hipLaunchKernelGGL(k_list_count, dim3(grid_of(np)), dim3(kTB), 0, s, np, pairs, geom, off, nchild, first_child,
                     theta, nearf, m2lc, nextc);
}

void list_write(hipStream_t s, int np, const uint64_t* pairs, const BoxGeom* geom, const double* off,
                const int32_t* nchild, const int32_t* first_child, double theta, const int32_t* nearf,
                const int32_t* near_at, const int32_t* m2l_at, const int32_t* next_at, uint64_t* near_out,
                uint64_t* m2l_out, uint64_t* next_out) {
  hipLaunchKernelGGL(k_list_write, dim3(grid_of(np)), dim3(kTB), 0, s, np, pairs, geom, off, nchild, first_child,
                     theta, nearf, near_at, m2l_at, next_at, near_out, m2l_out, next_out);
}

void list_totals(hipStream_t s, int np, const int32_t* nearf, const int32_t* m2lc, const int32_t* nextc,
                 const int32_t* near_at, const int32_t* m2l_at, const int32_t* next_at, int64_t* out) {
  hipLaunchKernelGGL(k_list_totals, dim3(1), dim3(64), 0, s, np, nearf, m2lc, nextc, near_at, m2l_at, next_at, out);
}

void csr_from_keys(hipStream_t s, int nb, const uint64_t* keys, int64_t n, int64_t* ptr, int32_t* idx) {
  const int64_t m = std::max<int64_t>(n, (int64_t)nb + 1);
  hipLaunchKernelGGL(k_csr_from_keys, dim3(grid_of(m)), dim3(kTB), 0, s, nb, keys, n, ptr, idx);
}

void node_keys(hipStream_t s, int N, const int32_t* target, uint64_t* keys) {
  if (N > 0) hipLaunchKernelGGL(k_node_keys, dim3(grid_of(N)), dim3(kTB), 0, s, N, target, keys);
}
void ns_tbox_flags(hipStream_t s, int nb, const int64_t* tptr, int32_t* flags) {
  hipLaunchKernelGGL(k_ns_tbox_flags, dim3(grid_of(nb)), dim3(kTB), 0, s, nb, tptr, flags);
}
void ns_tbox_compact(hipStream_t s, int nb, const int32_t* flags, const int32_t* at, int32_t* tbox) {
  hipLaunchKernelGGL(k_ns_tbox_compact, dim3(grid_of(nb)), dim3(kTB), 0, s, nb, flags, at, tbox);
}
void ns_chunk_count(hipStream_t s, int ntb, const int64_t* tptr, const int32_t* tbox, int32_t* cnt) {
  hipLaunchKernelGGL(k_ns_chunk_count, dim3(grid_of(ntb + 1)), dim3(kTB), 0, s, ntb, tptr, tbox, cnt);
}
void ns_range_count(hipStream_t s, int ntb, const int32_t* tbox, const int32_t* parent, const int64_t* near_ptr,
                    const int32_t* near_idx, const int32_t* nchild, int32_t* cnt) {
  hipLaunchKernelGGL(k_ns_range_count, dim3(grid_of(ntb + 1)), dim3(kTB), 0, s, ntb, tbox, parent, near_ptr, near_idx,
                     nchild, cnt);
}
void ns_range_write(hipStream_t s, int ntb, const int32_t* tbox, const int32_t* parent, const int64_t* near_ptr,
                    const int32_t* near_idx, const int32_t* nchild, const int64_t* box_range, const int32_t* at,
                    uint64_t* keys, int32_t* ends) {
  if (ntb > 0)
    hipLaunchKernelGGL(k_ns_range_write, dim3(grid_of(ntb)), dim3(kTB), 0, s, ntb, tbox, parent, near_ptr, near_idx,
                       nchild, box_range, at, keys, ends);
}
void ns_range_merge(hipStream_t s, int ntb, const int32_t* at, const int32_t* cnt, const uint64_t* keys,
                    const int32_t* ends, int64_t* mstart, int64_t* mend, int32_t* mcnt, int64_t* nfaces, int32_t* err) {
  hipLaunchKernelGGL(k_ns_range_merge, dim3(grid_of(ntb + 1)), dim3(kTB), 0, s, ntb, at, cnt, keys, ends, mstart, mend,
                     mcnt, nfaces, err);
}
void ns_dof_keys(hipStream_t s, int npf, int64_t total, int ntb, const int64_t* fstart, const int32_t* at,
                 const int32_t* mcnt, const int64_t* mstart, const int64_t* mend, const int32_t* fdof, uint64_t* keys) {
  if (total > 0)
    hipLaunchKernelGGL(k_ns_dof_keys, dim3(grid_of(total)), dim3(kTB), 0, s, npf, total, ntb, fstart, at, mcnt, mstart,
                       mend, fdof, keys);
}
void ns_face_cols(hipStream_t s, int npf, int64_t total, int ntb, const int64_t* fstart, const int32_t* at,
                  const int32_t* mcnt, const int64_t* mstart, const int64_t* mend, const int32_t* fdof,
                  const int64_t* dptr, const int32_t* dofs, int32_t* cols) {
  if (total > 0)
    hipLaunchKernelGGL(k_ns_face_cols, dim3(grid_of(total)), dim3(kTB), 0, s, npf, total, ntb, fstart, at, mcnt,
                       mstart, mend, fdof, dptr, dofs, cols);
}
void unique_flags(hipStream_t s, int64_t n, const uint64_t* keys, int32_t* flags) {
  if (n > 0) hipLaunchKernelGGL(k_unique_flags, dim3(grid_of(n)), dim3(kTB), 0, s, n, keys, flags);
}
void compact_u64(hipStream_t s, int64_t n, const uint64_t* keys, const int32_t* flags, const int32_t* at,
                 uint64_t* out) {
  if (n > 0) hipLaunchKernelGGL(k_compact_u64, dim3(grid_of(n)), dim3(kTB), 0, s, n, keys, flags, at, out);
}
void ns_chunk_keys(hipStream_t s, int ntb, const int64_t* tptr, const int32_t* tbox, const int32_t* chunk_at,
                   const int64_t* dptr, int C, uint64_t* keys, int32_t* err) {
  if (ntb > 0)
    hipLaunchKernelGGL(k_ns_chunk_keys, dim3(grid_of(ntb)), dim3(kTB), 0, s, ntb, tptr, tbox, chunk_at, dptr, C, keys,
                       err);
}
void ns_chunk_ti(hipStream_t s, int ntb, const int32_t* chunk_at, int32_t* chunk_ti) {
  if (ntb > 0) hipLaunchKernelGGL(k_ns_chunk_ti, dim3(grid_of(ntb)), dim3(kTB), 0, s, ntb, chunk_at, chunk_ti);
}
void ns_chunk_sizes(hipStream_t s, int nch, const uint64_t* order, const int32_t* chunk_ti, const int32_t* chunk_at,
                    const int64_t* tptr, const int32_t* tbox, const int64_t* dptr, const int32_t* mcnt, int C,
                    int64_t* rows_v, int64_t* cols_v, int64_t* mat_v, int64_t* rng_v, int64_t* nnz_v) {
  hipLaunchKernelGGL(k_ns_chunk_sizes, dim3(grid_of(nch + 1)), dim3(kTB), 0, s, nch, order, chunk_ti, chunk_at, tptr,
                     tbox, dptr, mcnt, C, rows_v, cols_v, mat_v, rng_v, nnz_v);
}
void ns_fill(hipStream_t s, int nch, const uint64_t* order, const int32_t* chunk_ti, const int32_t* chunk_at,
             const int64_t* tptr, const int32_t* tnodes, const int32_t* tbox, const int64_t* dptr, const int32_t* dofs,
             const int32_t* rat, const int32_t* mcnt, const int64_t* mstart, const int64_t* mend, int C, int n_dof,
             const int64_t* row_ptr, const int64_t* col_ptr, const int64_t* range_ptr, int32_t* row_out,
             int32_t* row_block, int64_t* node_row, int32_t* col_idx, int32_t* ncol_dof, int64_t* ranges,
             int64_t* range_prefix, const int64_t* fstart, int64_t* block_fstart) {
  if (nch > 0)
    hipLaunchKernelGGL(k_ns_fill, dim3(nch), dim3(kTB), 0, s, nch, order, chunk_ti, chunk_at, tptr, tnodes, tbox, dptr,
                       dofs, rat, mcnt, mstart, mend, C, n_dof, row_ptr, col_ptr, range_ptr, row_out, row_block,
                       node_row, col_idx, ncol_dof, ranges, range_prefix, fstart, block_fstart);
}
void l2p_coef(hipStream_t s, int64_t nrows, int P, const int32_t* row_out, const int32_t* target, const double* X,
              const BoxGeom* geom, const int32_t* level, const double* level_side, const double* cnm, double* out) {
  const int64_t n = nrows * (P + 1);
  if (n > 0)
    hipLaunchKernelGGL(k_l2p_coef, dim3(grid_of(n)), dim3(kTB), 0, s, nrows, P, row_out, target, X, geom, level,
                       level_side, cnm, out);
}
hipError_t scan_exclusive_i64(hipStream_t s, void* tmp, size_t& tmp_bytes, const int64_t* in, int64_t* out,
                              int64_t n) {
  return rocprim::exclusive_scan(tmp, tmp_bytes, in, out, (int64_t)0, (size_t)n, rocprim::plus<int64_t>(), s);
}

hipError_t sort_keys_u64(hipStream_t s, void* tmp, size_t& tmp_bytes, const uint64_t* in, uint64_t* out, int64_t n,
                         int begin_bit, int end_bit) {
  return rocprim::radix_sort_keys(tmp, tmp_bytes, in, out, (size_t)n, (unsigned)begin_bit, (unsigned)end_bit, s);
}

hipError_t sort_pairs_u64_i32(hipStream_t s, void* tmp, size_t& tmp_bytes, const uint64_t* keys_in,
                              uint64_t* keys_out, const int32_t* vals_in, int32_t* vals_out, int64_t n, int begin_bit,
                              int end_bit) {
  return rocprim::radix_sort_pairs(tmp, tmp_bytes, keys_in, keys_out, vals_in, vals_out, (size_t)n,
                                   (unsigned)begin_bit, (unsigned)end_bit, s);
}

hipError_t scan_inclusive_i32(hipStream_t s, void* tmp, size_t& tmp_bytes, const int32_t* in, int32_t* out,
                              int64_t n) {
  return rocprim::inclusive_scan(tmp, tmp_bytes, in, out, (size_t)n, rocprim::plus<int32_t>(), s);
}

hipError_t scan_exclusive_i32(hipStream_t s, void* tmp, size_t& tmp_bytes, const int32_t* in, int32_t* out,
                              int64_t n) {
  return rocprim::exclusive_scan(tmp, tmp_bytes, in, out, (int32_t)0, (size_t)n, rocprim::plus<int32_t>(), s);
}

}  // namespace dev
}  // namespace bemfmm
