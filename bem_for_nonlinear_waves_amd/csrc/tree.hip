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
  if (i >= n) return;
  const int f = (int)(uint32_t)keys[i];
  const int cell = (int)((uint32_t)(keys[i] >> 32) % 32u);
  const int c = le + ci[i] - 1;
  const double x[3] = {pole[3 * f], pole[3 * f + 1], pole[3 * f + 2]};
  if (box_inside(geom[c], x)) {
    fbox[f] = c;
    fkey[f] |= (uint64_t)(cell + 1);
    atomicAdd(&count[c], 1);
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
