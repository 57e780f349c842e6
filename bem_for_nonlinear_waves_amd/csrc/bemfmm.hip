// MI355X (gfx950) device side of the FMM-accelerated GMRES BVP engine and its C ABI.
//
// Matvec pipeline on one HIP stream, all data resident in HBM (the reference's
// compute_Ax_minus_b, bem/BEM_solveBVP_GMRES_FMM.hpp:34-158, with updateFMM,
// lib/include/moments.hpp:1547-1574):
//   scatter   x + knowns -> phi_n, phi per DOF                (setUnknowns/setKnowns, :1142-1176)
//   P2M       per-leaf dense GEMV over the leaf's DOF columns (increment_M_reuse, moments.hpp:358-368)
//   M2M/M2L/L2L  operator-grouped batched dense products       (moments.hpp:419-436, 732-764, 869-884)
//   near+L2P  per-target-box dense GEMV over near DOF columns and the box's local coefficients
//             (integrateNearField/integrateFarField, lib_multipole_expansion.hpp:1143-1150, 1273-1302)
//   assemble  Av_i = GPhin - (GnPhi + alpha_i phi_i); CORNER rows Av_i = phi_i
// Setup: the host builds the plan (plan.cpp); the near-field quadrature
// (GW5x5 singular/near, Dunavant far; GMRES_FMM.hpp:1824-1891) runs on the GPU.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <utility>
#include <array>
#include <chrono>
#include <map>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/bemfmm.h"
#include "nearquad.hpp"
#include "plan.hpp"

namespace {

using bemfmm::Plan;

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

// ---------------------------------------------------------------------------
// GW5x5 product rule on the unit square, (t0, t1, w) (reference basic_constants.hpp:1326-1350).
// ---------------------------------------------------------------------------
__constant__ double c_gw5[25][3] = {
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

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ void k_fill(double* __restrict__ p, int64_t n, double v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

// mode 0: matvec (unknowns = x, knowns = 0); 1: rhs (unknowns = 0, knowns = BC); 2: rigid mode (phi = 1, phi_n = 0).
// A ROBIN DOF (kind 3) has the unknown phi and phi_n = kappa phi + g: kappa phi belongs to the matvec, the
// known g (passed as its phi_n) to the right-hand side (main_freq_domain.cpp:543-555 with g = 0).
__global__ void k_scatter(int mode, int n, const int32_t* __restrict__ kind, const double* __restrict__ kappa,
                          const double* __restrict__ x, const double* __restrict__ phi_k,
                          const double* __restrict__ phin_k, double* __restrict__ phin, double* __restrict__ phi) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool dir = kind[i] == 0;
  double a = 0., b = 0.;  // a = phi_n, b = phi
  if (mode == 0) {
    if (dir) {
      a = x[i];
    } else {
      b = x[i];
      if (kind[i] == 3) a = kappa[i] * x[i];
    }
  } else if (mode == 1) {
    if (dir) b = phi_k[i];
    else a = phin_k[i];
  } else {
    b = 1.0;
  }
  phin[i] = a;
  phi[i] = b;
}

// Box-blocked dense GEMV (see bemfmm::BlockGemv). One workgroup per block; the gathered
// inputs are staged through LDS in chunks; each wavefront streams kRPW rows at once with
// 16-byte vector loads (double2 / float4: rows are padded to a multiple of 4 columns and start
// 16-byte aligned) and issues all 2*kRPW*2 loads of two column steps before using them (rows
// past the end of the block are clamped to its last row and their sums discarded), so that
// enough HBM requests are in flight per lane. Sums are accumulated in FP64.
// part 0: all columns, y = A x. The columns of a block are [DOF columns (ncd) | L2P columns | padding];
// part 1: the DOF columns [0, ncd rounded down to 4), y = A x (needs no FMM result: it runs beside the
// FMM passes); part 2: the rest (0-3 DOF columns and the L2P columns), y += A x (after L2L).
// Streaming (non-temporal) loads of data read once per matvec (the near-field matrices, the staged M2L
// results): they are far larger than L2 and would only evict what is reused (gathered vectors,
// multipoles, operators).
__device__ __forceinline__ float4 ld_nt(const float4* p) {
  typedef float v4 __attribute__((ext_vector_type(4)));
  const v4 r = __builtin_nontemporal_load(reinterpret_cast<const v4*>(p));
  return make_float4(r.x, r.y, r.z, r.w);
}
__device__ __forceinline__ double2 ld_nt(const double2* p) {
  typedef double v2 __attribute__((ext_vector_type(2)));
  const v2 r = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
  return make_double2(r.x, r.y);
}
__device__ __forceinline__ float2 ld_nt(const float2* p) {
  typedef float v2 __attribute__((ext_vector_type(2)));
  const v2 r = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
  return make_float2(r.x, r.y);
}
constexpr int kGemvChunk = 2048;
constexpr int kGemvMaxRows = 512;
constexpr int kRPW = 4;
template <typename T>
struct GemvVec;
template <>
struct GemvVec<double> {
  typedef double2 type;
  static constexpr int W = 2;
  static __device__ __forceinline__ double get(const type& v, int i) { return i == 0 ? v.x : v.y; }
};
template <>
struct GemvVec<float> {
  typedef float4 type;
  static constexpr int W = 4;
  static __device__ __forceinline__ double get(const type& v, int i) {
    return (double)(i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w);
  }
};
template <typename T>
__global__ __launch_bounds__(256) void k_block_gemv(const int64_t* __restrict__ row_ptr, const int64_t* __restrict__ col_ptr,
                                                    const int64_t* __restrict__ mat_ptr, const int32_t* __restrict__ row_out,
                                                    const int32_t* __restrict__ col_idx, const T* __restrict__ mat,
                                                    const double* __restrict__ x0, const double* __restrict__ x1,
                                                    double* __restrict__ y0, double* __restrict__ y1, int part,
                                                    const int32_t* __restrict__ ncol_dof) {
  typedef typename GemvVec<T>::type V;
  constexpr int W = GemvVec<T>::W;
  constexpr int STEP = 64 * W;  // columns covered by one wave-wide vector load
  __shared__ double xs0[kGemvChunk], xs1[kGemvChunk];
  __shared__ double acc[2 * kGemvMaxRows];
  const int b = blockIdx.x;
  const int64_t r0 = row_ptr[b];
  const int nrow = (int)(row_ptr[b + 1] - r0);
  const int64_t c0 = col_ptr[b];
  const int ncol = (int)(col_ptr[b + 1] - c0);  // multiple of 4
  const T* M = mat + mat_ptr[b];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < 2 * nrow; i += blockDim.x) acc[i] = 0.;
  const int split = part == 0 ? 0 : (ncol_dof[b] & ~3);
  const int cb = part == 2 ? split : 0, ce = part == 1 ? split : ncol;
  for (int cs = cb; cs < ce; cs += kGemvChunk) {
    const int cn = min(kGemvChunk, ce - cs);  // multiple of 4
    __syncthreads();
    const int32_t* ci = col_idx + c0 + cs;
    int i = threadIdx.x;
    for (; i + 3 * 256 < cn; i += 4 * 256) {
      const int g0 = ci[i], g1 = ci[i + 256], g2 = ci[i + 512], g3 = ci[i + 768];
      const double a0 = x0[g0], a1 = x0[g1], a2 = x0[g2], a3 = x0[g3];
      const double b0 = x1[g0], b1 = x1[g1], b2 = x1[g2], b3 = x1[g3];
      xs0[i] = a0;
      xs0[i + 256] = a1;
      xs0[i + 512] = a2;
      xs0[i + 768] = a3;
      xs1[i] = b0;
      xs1[i + 256] = b1;
      xs1[i + 512] = b2;
      xs1[i + 768] = b3;
    }
    for (; i < cn; i += 256) {
      const int g = ci[i];
      xs0[i] = x0[g];
      xs1[i] = x1[g];
    }
    __syncthreads();
    for (int rg = wave * kRPW; rg < nrow; rg += 4 * kRPW) {
      const T* Ar[kRPW];
#pragma unroll
      for (int k = 0; k < kRPW; ++k) Ar[k] = M + (int64_t)min(rg + k, nrow - 1) * 2 * ncol + cs;
      double s0[kRPW], s1[kRPW];
#pragma unroll
      for (int k = 0; k < kRPW; ++k) s0[k] = s1[k] = 0.;
      int c = lane * W;
      for (; c + STEP < cn; c += 2 * STEP) {
        V va[kRPW][2], vb[kRPW][2];
#pragma unroll
        for (int k = 0; k < kRPW; ++k) {
          va[k][0] = ld_nt(reinterpret_cast<const V*>(Ar[k] + c));
          va[k][1] = ld_nt(reinterpret_cast<const V*>(Ar[k] + c + STEP));
          vb[k][0] = ld_nt(reinterpret_cast<const V*>(Ar[k] + ncol + c));
          vb[k][1] = ld_nt(reinterpret_cast<const V*>(Ar[k] + ncol + c + STEP));
        }
        double xa[2][W], xb[2][W];
#pragma unroll
        for (int u = 0; u < W; ++u) {
          xa[0][u] = xs0[c + u];
          xa[1][u] = xs0[c + STEP + u];
          xb[0][u] = xs1[c + u];
          xb[1][u] = xs1[c + STEP + u];
        }
#pragma unroll
        for (int k = 0; k < kRPW; ++k)
#pragma unroll
          for (int u = 0; u < W; ++u) {
            s0[k] += GemvVec<T>::get(va[k][0], u) * xa[0][u] + GemvVec<T>::get(va[k][1], u) * xa[1][u];
            s1[k] += GemvVec<T>::get(vb[k][0], u) * xb[0][u] + GemvVec<T>::get(vb[k][1], u) * xb[1][u];
          }
      }
      if (c < cn) {
        V va[kRPW], vb[kRPW];
#pragma unroll
        for (int k = 0; k < kRPW; ++k) {
          va[k] = ld_nt(reinterpret_cast<const V*>(Ar[k] + c));
          vb[k] = ld_nt(reinterpret_cast<const V*>(Ar[k] + ncol + c));
        }
#pragma unroll
        for (int k = 0; k < kRPW; ++k)
#pragma unroll
          for (int u = 0; u < W; ++u) {
            s0[k] += GemvVec<T>::get(va[k], u) * xs0[c + u];
            s1[k] += GemvVec<T>::get(vb[k], u) * xs1[c + u];
          }
      }
#pragma unroll
      for (int k = 0; k < kRPW; ++k) {
        double u = s0[k], v = s1[k];
        for (int off = 32; off > 0; off >>= 1) {
          u += __shfl_down(u, off);
          v += __shfl_down(v, off);
        }
        if (lane == 0 && rg + k < nrow) {
          acc[2 * (rg + k)] += u;
          acc[2 * (rg + k) + 1] += v;
        }
      }
    }
  }
  __syncthreads();
  for (int r = threadIdx.x; r < nrow; r += blockDim.x) {
    const int o = row_out[r0 + r];
    if (part == 2) {  // each row belongs to one block: a plain read-modify-write
      y0[o] += acc[2 * r];
      y1[o] += acc[2 * r + 1];
    } else {
      y0[o] = acc[2 * r];
      y1[o] = acc[2 * r + 1];
    }
  }
}

// P2M: one workgroup per leaf; the leaf's matrix is column-major per layer (plan.hpp), so thread
// (layer = tid / 128, row = tid % 128 + 128 k) streams a column's rows with its neighbours: every load is
// coalesced and every lane has a row, however few DOF columns the leaf has. The gathered phi_n / phi
// values are staged in LDS in chunks. M[leaf][row] = sum_col A[layer][col][row] x_layer[col] (FP64 sums).
constexpr int kP2MChunk = 1024;
template <typename T>
__global__ __launch_bounds__(256) void k_p2m(int C, int CP, const int64_t* __restrict__ row_ptr,
                                             const int64_t* __restrict__ col_ptr, const int64_t* __restrict__ mat_ptr,
                                             const int32_t* __restrict__ row_out, const int32_t* __restrict__ col_idx,
                                             const T* __restrict__ mat, const double* __restrict__ x0,
                                             const double* __restrict__ x1, double* __restrict__ M0,
                                             double* __restrict__ M1) {
  __shared__ double xs[2][kP2MChunk];
  constexpr int kRows = 4;  // rows per thread: C <= 512
  const int b = blockIdx.x;
  const int64_t c0 = col_ptr[b];
  const int ncol = (int)(col_ptr[b + 1] - c0);
  const T* A = mat + mat_ptr[b];
  const int layer = threadIdx.x >> 7, r0 = threadIdx.x & 127;
  const T* Al = A + (int64_t)layer * ncol * CP;
  double acc[kRows] = {0., 0., 0., 0.};
  for (int cs = 0; cs < ncol; cs += kP2MChunk) {
    const int cn = min(kP2MChunk, ncol - cs);
    __syncthreads();
    for (int i = threadIdx.x; i < cn; i += blockDim.x) {
      const int g = col_idx[c0 + cs + i];
      xs[0][i] = x0[g];
      xs[1][i] = x1[g];
    }
    __syncthreads();
    const double* xl = xs[layer];
#pragma unroll
    for (int k = 0; k < kRows; ++k) {
      const int r = r0 + 128 * k;
      if (r >= CP) break;
      const T* a = Al + (int64_t)cs * CP + r;
      double s0 = 0., s1 = 0.;
      int c = 0;
      for (; c + 7 < cn; c += 8) {  // 8 loads in flight (streamed: non-temporal)
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = (double)__builtin_nontemporal_load(&a[(int64_t)(c + u) * CP]);
#pragma unroll
        for (int u = 0; u < 8; u += 2) {
          s0 += v[u] * xl[c + u];
          s1 += v[u + 1] * xl[c + u + 1];
        }
      }
      for (; c < cn; ++c) s0 += (double)__builtin_nontemporal_load(&a[(int64_t)c * CP]) * xl[c];
      acc[k] += s0 + s1;
    }
  }
  const int64_t base = row_out[row_ptr[b]];  // leaf box * C (rows of a leaf are consecutive)
  double* out = layer ? M1 : M0;
#pragma unroll
  for (int k = 0; k < kRows; ++k) {
    const int r = r0 + 128 * k;
    if (r < C) out[base + r] = acc[k];
  }
}

// Operator-grouped translation on the matrix cores. For every pair (dst, src) of a
// tile (<= TP pairs sharing one operator): D[dst] += T_op * S[src] for both layers.
// The tile is the GEMM  out[RP x 2TP] += T[RP x KP] * B[KP x 2TP]  built from
// 16x16x4 MFMA blocks (RP = 64*RTW rows: C padded, KP = C padded to 16, zeros in the
// padding): wave w owns row tiles [w*RTW, (w+1)*RTW) and all 2TP/16 16-column tiles.
// B (the gathered source coefficients, rounded to T) is staged in LDS; T is streamed
// from the padded operator array [op][KP][RP] through a two-stage register pipeline of
// kKG k-steps, so that the loads of the next stage are in flight while the current one runs.
// Workgroups are remapped so that each XCD processes one contiguous run of tiles
// (tiles are sorted by operator): the tiles of an operator share that XCD's L2.
// T = double: v_mfma_f64_16x16x4f64; T = float: v_mfma_f32_16x16x4f32 (FP32 accumulation).
constexpr int kTP = 8;       // pairs per tile of the M2M/L2L translations: their small upper levels are
                             // latency-bound, so more, smaller workgroups pay (209k tank, per matvec:
                             // M2M and L2L each 0.111 ms at 32 pairs, 0.091 at 16, 0.078 at 8)
constexpr int kTPM2L = 16;   // default pairs per M2L tile: 3 workgroups per CU (FP64) instead of 2 at 32 pairs;
                             // measured on the 209k tank: FP64 3.61 -> 3.41 ms, FP32 2.52 -> 2.17 ms; 64 pairs: 4.6 ms
// LDS elements of a tile of TP pairs: the B stage (KP x (2TP+1)) or the output transpose (2TP x (RP+1))
inline size_t translate_lds(int TP, int KP, int RP) { return (size_t)std::max(KP * (2 * TP + 1), 2 * TP * (RP + 1)); }
constexpr int kXCD = 8;
constexpr int kKG = 4;  // k-steps (of 4) per pipeline stage
typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef double4_t acc_t;
  static __device__ __forceinline__ acc_t run(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // D block layout (measured on gfx950): register i of lane l holds row l/16 + 4*i of column l%16
  static __device__ __forceinline__ int row(int lk, int i) { return lk + 4 * i; }
};
template <>
struct Mfma<float> {
  typedef float4_t acc_t;
  static __device__ __forceinline__ acc_t run(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // D block layout (measured on gfx950): register i of lane l holds row 4*(l/16) + i of column l%16
  static __device__ __forceinline__ int row(int lk, int i) { return 4 * lk + i; }
};

// Values per staged M2L column. FP32: C rounded up to 32, so every column starts on a 128-byte line and the
// staging is written and read in whole lines (at stride C = 121 the rotation M2L's 16-row stores straddled
// lines, and its writes and the reduction's reads moved 14% and 7% more than the staged bytes); C <= 128
// (checked in prepare) keeps it <= 128. FP64: C (each wave writes whole 968-byte columns at once: no
// partial lines to remove, and the padding only added bytes).
template <typename T>
__host__ __device__ constexpr int stage_stride(int C) { return sizeof(T) == 4 ? (C + 31) & ~31 : C; }
// STAGE = false: atomically accumulate into D0/D1 at the pair's dst box (M2M, L2L).
// STAGE = true: write each pair's two result columns to stage[(pair*2 + layer)*CS + row] (M2L), CS =
// stage_stride<T>(C) (rows C..CS-1 written as zeros).
template <typename T, int RTW, bool STAGE, int TP>
__global__ __launch_bounds__(256) void k_translate_mfma(int C, int RP, int KP, int ntiles, const T* __restrict__ ops,
                                                        const int32_t* __restrict__ tile_op, const int64_t* __restrict__ tile_p0,
                                                        const int64_t* __restrict__ tile_p1, const int32_t* __restrict__ dst,
                                                        const int32_t* __restrict__ src, const double* __restrict__ S0,
                                                        const double* __restrict__ S1, double* __restrict__ D0,
                                                        double* __restrict__ D1, T* __restrict__ stage) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int kTCols = 2 * TP;
  constexpr int kTStride = kTCols + 1;
  constexpr int kTCT = kTCols / 16;  // 16-column MFMA tiles per workgroup tile
  T* Bs = reinterpret_cast<T*>(smem);  // KP * kTStride
  const int per_xcd = (ntiles + kXCD - 1) / kXCD;
  const int t = (blockIdx.x % kXCD) * per_xcd + blockIdx.x / kXCD;
  if (t >= ntiles) return;
  const int op = tile_op[t];
  const int64_t p0 = tile_p0[t];
  const int np = (int)(tile_p1[t] - p0);
  {
    // gather the tile's source coefficients: thread -> (layer, k); 32 loads issued together
    const int layer = threadIdx.x >> 7;
    const double* S = layer ? S1 : S0;
    constexpr int G = TP < 32 ? TP : 32;
#pragma unroll
    for (int u0 = 0; u0 < TP; u0 += G) {
      int sidx[G];
#pragma unroll
      for (int u = 0; u < G; ++u) sidx[u] = (u0 + u < np) ? src[p0 + u0 + u] : -1;
      for (int k = threadIdx.x & 127; k < KP; k += 128) {
        double v[G];
#pragma unroll
        for (int u = 0; u < G; ++u) v[u] = (sidx[u] >= 0 && k < C) ? S[(int64_t)sidx[u] * C + k] : 0.;
#pragma unroll
        for (int u = 0; u < G; ++u) Bs[k * kTStride + 2 * (u0 + u) + layer] = (T)v[u];
      }
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const T* O = ops + (int64_t)op * KP * RP + (int64_t)lk * RP + wave * RTW * 16 + li;
  const T* Bl = Bs + lk * kTStride + li;
  typename Mfma<T>::acc_t acc[RTW][kTCT];
#pragma unroll
  for (int a = 0; a < RTW; ++a)
#pragma unroll
    for (int ct = 0; ct < kTCT; ++ct) acc[a][ct] = typename Mfma<T>::acc_t{0, 0, 0, 0};
  T av[2][kKG][RTW];
  T bv[2][kKG][kTCT];
  auto load = [&](int st, int k0) {
#pragma unroll
    for (int s = 0; s < kKG; ++s) {
      const T* ok = O + (int64_t)(k0 + 4 * s) * RP;
#pragma unroll
      for (int a = 0; a < RTW; ++a) av[st][s][a] = ok[a * 16];
      const T* bk = Bl + (k0 + 4 * s) * kTStride;
#pragma unroll
      for (int ct = 0; ct < kTCT; ++ct) bv[st][s][ct] = bk[16 * ct];
    }
  };
  auto run = [&](int st) {
#pragma unroll
    for (int s = 0; s < kKG; ++s)
#pragma unroll
      for (int a = 0; a < RTW; ++a)
#pragma unroll
        for (int ct = 0; ct < kTCT; ++ct) acc[a][ct] = Mfma<T>::run(av[st][s][a], bv[st][s][ct], acc[a][ct]);
  };
  const int ng = KP / (4 * kKG);
  load(0, 0);
  for (int g = 0; g < ng; g += 2) {
    if (g + 1 < ng) load(1, (g + 1) * 4 * kKG);
    run(0);
    if (g + 1 < ng) {
      if (g + 2 < ng) load(0, (g + 2) * 4 * kKG);
      run(1);
    }
  }
  if (STAGE) {
    // per-pair results -> staging rows with coalesced stores: transpose the tile through
    // LDS (column-major, stride RP+1) and let each wave write whole columns
    __syncthreads();
    T* Tt = Bs;  // reuse the staging area (sized by translate_lds)
#pragma unroll
    for (int a = 0; a < RTW; ++a) {
      const int rt = wave * RTW + a;
#pragma unroll
      for (int ct = 0; ct < kTCT; ++ct)
#pragma unroll
        for (int i = 0; i < 4; ++i) Tt[(ct * 16 + li) * (RP + 1) + rt * 16 + Mfma<T>::row(lk, i)] = acc[a][ct][i];
    }
    __syncthreads();
    for (int col = wave; col < 2 * np; col += 4) {
      const int pair = col >> 1, layer = col & 1;
      const int CS = stage_stride<T>(C);  // <= RP
      T* out = stage + ((p0 + pair) * 2 + layer) * (int64_t)CS;
      // FP64 staging is stored non-temporally: its reduction then ran 0.85 -> 0.81 ms at 209k nodes (the
      // FP32 rotation M2L's stores stay ordinary: non-temporal ones made it 14% slower)
      for (int r = lane; r < CS; r += 64) {
        const T val = r < C ? Tt[col * (RP + 1) + r] : T(0);
        if constexpr (sizeof(T) == 8) __builtin_nontemporal_store(val, &out[r]);
        else out[r] = val;
      }
    }
    return;
  }
#pragma unroll
  for (int a = 0; a < RTW; ++a) {
    const int rt = wave * RTW + a;
#pragma unroll
    for (int ct = 0; ct < kTCT; ++ct) {
      const int col = ct * 16 + li;
      const int pair = col >> 1, layer = col & 1;
      if (pair >= np) continue;
      double* D = (layer ? D1 : D0) + (int64_t)dst[p0 + pair] * C;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = rt * 16 + Mfma<T>::row(lk, i);
        if (row < C) atomicAdd(&D[row], (double)acc[a][ct][i]);
      }
    }
  }
}

// M2L reduction: L[t] = sum over the M2L pairs of box t (fixed order) of the staged per-pair
// results (FP64 accumulation). One workgroup per box, the boxes launched in descending pair count
// (order: the longest reductions start first, no straggler tail); thread j < 2C owns (layer j/C, row
// j%C). Writes every box, so L needs no clearing before it.
template <typename T>
__global__ __launch_bounds__(256) void k_m2l_reduce(int C, const int32_t* __restrict__ order,
                                                    const int64_t* __restrict__ tptr, const int32_t* __restrict__ tpairs,
                                                    const T* __restrict__ stage, double* __restrict__ L0,
                                                    double* __restrict__ L1) {
  const int t = order[blockIdx.x];
  const int j = threadIdx.x;
  if (j >= 2 * C) return;
  const int64_t a = tptr[t], b = tptr[t + 1];
  const int CS = stage_stride<T>(C);
  const int jj = j < C ? j : CS + j - C;  // (layer, row) -> offset in the pair's two staged columns
  double s[4] = {0., 0., 0., 0.};
  int64_t q = a;
  for (; q + 7 < b; q += 8) {  // 8 loads in flight
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = __builtin_nontemporal_load(&stage[(int64_t)tpairs[q + u] * 2 * CS + jj]);
#pragma unroll
    for (int u = 0; u < 8; ++u) s[u & 3] += v[u];
  }
  for (; q < b; ++q) s[0] += __builtin_nontemporal_load(&stage[(int64_t)tpairs[q] * 2 * CS + jj]);
  const double v = (s[0] + s[1]) + (s[2] + s[3]);
  if (j < C) L0[(int64_t)t * C + j] = v;
  else L1[(int64_t)t * C + j - C] = v;
}

// M2L by rotation, FP32 operator: per pair L = H^T T(|o| z) D M, where D and H^T are the per-degree
// ((2n+1)^2) real rotation blocks taking the offset o onto the z axis and T(|o| z) is the axial operator
// (one (P+1-m)^2 block per order m, the same for the real and imaginary parts), with the z-rotations as
// 2x2 blocks per order: 3113 multiply-adds per column at P = 10 instead of 14641. A workgroup takes up to
// 128 pairs of one offset: the factors are loaded once into LDS, and each factor row is spread over the
// lanes of every 16-lane row and broadcast to the multiply-adds through DPP (rot_dot); each of the 4 waves
// owns 32 pairs, lane 2u + layer holding column (pair u, layer) in registers. Columns move between the
// [box][C] layouts and the lanes through a per-wave LDS buffer, 16 rows at a time (coalesced loads and
// stores). Output: stage[(pair*2 + layer)*CS + row], as the matrix-core M2L with STAGE.
constexpr int kRotWavePairs = 32;
constexpr int kRotTP = 4 * kRotWavePairs;  // pairs per workgroup tile
// Device layout of the factors of one offset, in float4s, every row zero-padded to whole float4s
// (the z-rotation pairs are read as float4s; a factor row is read one float per lane, lane l taking
// entry l & 15): (cos m phi, sin m phi) for m = 1..P; per degree n the y-rotation blocks
// Y_D (cos-type (n+1) rows, sin-type n rows); per order m the axial block (P+1-m rows); per degree the
// blocks of Y_H^T. Row lengths equal the block sizes.
__host__ __device__ constexpr int rot_q(int len) { return (len + 3) / 4; }
__host__ __device__ constexpr int rot_size4(int P) {
  int s = rot_q(2 * P);
  for (int n = 0; n <= P; ++n) s += 2 * ((n + 1) * rot_q(n + 1) + n * rot_q(n));
  for (int m = 0; m <= P; ++m) s += (P + 1 - m) * rot_q(P + 1 - m);
  return s;
}
__host__ __device__ constexpr int rot_size(int P) { return 4 * rot_size4(P); }  // floats per offset
// compact factors (m2l_rotation_operator) -> the padded device layout
void rot_pad(int P, const double* in, float* out) {
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
__device__ constexpr int rot_re(int n, int m) { return m == 0 ? n * n : n * n + 2 * m - 1; }
__device__ constexpr int rot_im(int n, int m) { return n * n + 2 * m; }
// LDS writes and reads of one wave's lanes, ordered (the buffer is private to the wave)
__device__ __forceinline__ void wave_lds_sync() {
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  __builtin_amdgcn_wave_barrier();
}

// Rows of the three factor stages, in execution order: stage 0 = Y_D (degree n: cos-type rows a = 0..n,
// then sin-type rows a = 0..n-1), 1 = axial (order m, rows j = m + a), 2 = Y_H^T (as stage 0). Each row is
// a dot product of one padded factor row (q float4s at float4 offset off) with the lane's coefficients.
// The z-rotations (2x2 per order) run before a degree's Y_D rows and after its Y_H^T rows. The rows run
// software-pipelined: the factors of row r+D are read from LDS before the multiply-adds of row r, with
// scheduling barriers in between, so every LDS read has D rows of FMAs to complete.
struct RotRow {
  int stage, blk, kind, a, len, q, off;  // kind: 0 = cos-type, 1 = sin-type (stages 0, 2)
};
template <int P>
__host__ __device__ constexpr int rot_rows() { return 2 * (P + 1) * (P + 1) + (P + 1) * (P + 2) / 2; }
// descriptor of row K (compile-time)
template <int P>
__host__ __device__ constexpr RotRow rot_row(int K) {
  int k = 0, off = rot_q(2 * P);
  for (int st = 0; st < 3; ++st) {
    if (st == 1) {
      for (int m = 0; m <= P; ++m) {
        const int w = P + 1 - m;
        if (K < k + w) return RotRow{1, m, 0, K - k, w, rot_q(w), off + (K - k) * rot_q(w)};
        k += w;
        off += w * rot_q(w);
      }
      continue;
    }
    for (int n = 0; n <= P; ++n)
      for (int kind = 0; kind < 2; ++kind) {
        const int w = kind == 0 ? n + 1 : n;
        if (K < k + w) return RotRow{st, n, kind, K - k, w, rot_q(w), off + (K - k) * rot_q(w)};
        k += w;
        off += w * rot_q(w);
      }
  }
  return RotRow{-1, 0, 0, 0, 0, 0, 0};
}
// packed index of entry i of the cos-type (Re of m = 0..n) / sin-type (Im of m = 1..n) list of degree n
__device__ constexpr int rot_cs(int n, int kind, int i) { return kind == 0 ? rot_re(n, i) : rot_im(n, i + 1); }

// 0, but opaque to the compiler and available only once `dep` is computed: offsetting the reads of row
// r+1 by it (dep = the result of row r-1) keeps the compiler from clustering and hoisting the factor reads
// of later rows (with one base address and constant offsets it hoists thousands of them and spills)
__device__ __forceinline__ int rot_after(float dep) {
  int z;
  asm("v_mov_b32 %0, 0" : "=v"(z) : "v"(dep));
  return z;
}

// Factor broadcast across lanes: lane l holds entry (l & 15) of one factor row (one ds_read_b32 per row,
// the wave's four 16-lane rows read the same 16 addresses), and each multiply-add takes entry i from lane
// i of its 16-lane row through the DPP row_newbcast operand. The LDS returns 4 bytes per lane per factor
// row this way, instead of 16 bytes per lane per 4 entries for an all-lanes broadcast read, which made
// the LDS return path the bound. The compiler does not form DPP multiply-adds from ordinary code, so a
// row's dot product sum_i f[i] x[i] is one asm block (the result is early-clobber: it is written before
// the last inputs are read). s_nop 1: a DPP source written by a VALU instruction needs two wait states,
// which the hazard recognizer does not insert for inline asm (the factor register comes straight from
// an LDS load; this covers a copy the register allocator might place before the row). Inside the block
// only the accumulator is carried from one instruction to the next, which needs no wait states.
#define ROT_M0 "v_mul_f32_dpp %0, %1, %2 row_newbcast:0 row_mask:0xf bank_mask:0xf"
#define ROT_F(i, o) "\n\tv_fmac_f32_dpp %0, %1, %" #o " row_newbcast:" #i " row_mask:0xf bank_mask:0xf"
#define ROT_D1 ROT_M0
#define ROT_X1 "v"(x[0])
#define ROT_D2 ROT_D1 ROT_F(1, 3)
#define ROT_X2 ROT_X1, "v"(x[1])
#define ROT_D3 ROT_D2 ROT_F(2, 4)
#define ROT_X3 ROT_X2, "v"(x[2])
#define ROT_D4 ROT_D3 ROT_F(3, 5)
#define ROT_X4 ROT_X3, "v"(x[3])
#define ROT_D5 ROT_D4 ROT_F(4, 6)
#define ROT_X5 ROT_X4, "v"(x[4])
#define ROT_D6 ROT_D5 ROT_F(5, 7)
#define ROT_X6 ROT_X5, "v"(x[5])
#define ROT_D7 ROT_D6 ROT_F(6, 8)
#define ROT_X7 ROT_X6, "v"(x[6])
#define ROT_D8 ROT_D7 ROT_F(7, 9)
#define ROT_X8 ROT_X7, "v"(x[7])
#define ROT_D9 ROT_D8 ROT_F(8, 10)
#define ROT_X9 ROT_X8, "v"(x[8])
#define ROT_D10 ROT_D9 ROT_F(9, 11)
#define ROT_X10 ROT_X9, "v"(x[9])
#define ROT_D11 ROT_D10 ROT_F(10, 12)
#define ROT_X11 ROT_X10, "v"(x[10])
template <int L>
__device__ __forceinline__ float rot_dot(float f, const float* x);
#define ROT_DOT(L)                                                                  \
  template <>                                                                       \
  __device__ __forceinline__ float rot_dot<L>(float f, const float* x) {            \
    float r;                                                                        \
    asm("s_nop 1\n\t" ROT_D##L : "=&v"(r) : "v"(f), ROT_X##L);                       \
    return r;                                                                       \
  }
ROT_DOT(1)
ROT_DOT(2)
ROT_DOT(3)
ROT_DOT(4)
ROT_DOT(5)
ROT_DOT(6)
ROT_DOT(7)
ROT_DOT(8)
ROT_DOT(9)
ROT_DOT(10)
ROT_DOT(11)
#undef ROT_DOT

// y-rotation row of degree N: sum_i B[a][i] v[list(i)] over the degree's cos- (KIND 0) or sin-type list
template <int N, int KIND, int... I>
__device__ __forceinline__ float rot_ydot(float f, const float* v, std::integer_sequence<int, I...>) {
  const float x[] = {v[rot_cs(N, KIND, I)]...};
  return rot_dot<sizeof...(I)>(f, x);
}
// axial row of order M: sum over n = M + i of T[a][i] (Re, Im of (n, M))
template <int M, int... I>
__device__ __forceinline__ void rot_adot(float f, const float* v, float& sr, float& si, std::integer_sequence<int, I...>) {
  const float xr[] = {v[rot_re(M + I, M)]...};
  sr = rot_dot<sizeof...(I)>(f, xr);
  if constexpr (M > 0) {
    const float xi[] = {v[rot_im(M + I, M)]...};
    si = rot_dot<sizeof...(I)>(f, xi);
  }
}

typedef __attribute__((address_space(3))) const float rot_lds_float;
template <int P>
struct RotPipe {
  static constexpr int D = 4;  // factor rows read ahead
  rot_lds_float* R;            // the factors offset by (lane & 15) floats
  float f[D + 1];
  float tmp[P + 1], ti[P + 1];
  float last = 0.f;  // result of the previous row
};

// (Re, Im) of orders 1..n of degree n rotated about z: forward (c re - s im, s re + c im), or transposed
template <int P, bool kTransposed>
__device__ __forceinline__ void rot_z(const float4* __restrict__ R4, float* v, int n) {
#pragma unroll
  for (int m = 1; m <= n; ++m) {
    const float4 cs = R4[(m - 1) / 2];
    const float c = (m & 1) ? cs.x : cs.z, s = (m & 1) ? cs.y : cs.w;
    const float re = v[rot_re(n, m)], im = v[rot_im(n, m)];
    if (kTransposed) {
      v[rot_re(n, m)] = fmaf(c, re, s * im);
      v[rot_im(n, m)] = fmaf(c, im, -s * re);
      // final values: computed here, not sunk to the scatter (which would keep every degree's inputs live)
      asm volatile("" : "+v"(v[rot_re(n, m)]), "+v"(v[rot_im(n, m)]));
    } else {
      v[rot_re(n, m)] = fmaf(c, re, -s * im);
      v[rot_im(n, m)] = fmaf(c, im, s * re);
    }
  }
}

// Row K (R4 = the factors as float4s, for the z-rotation pairs). The read of row K + D is issued first;
// its base pointer passes through an empty asm that takes row K - 1's result as an input, so the read
// cannot be hoisted above that row (with one base and constant offsets the compiler hoists and clusters
// thousands of reads) and the row's offset stays the load's immediate.
template <int P, int K>
__device__ __forceinline__ void rot_step(const float4* __restrict__ R4, float* v, RotPipe<P>& s) {
  constexpr RotRow rw = rot_row<P>(K);
  constexpr int D = RotPipe<P>::D;
  if constexpr (K + D < rot_rows<P>()) {
    asm("" : "+v"(s.R) : "v"(s.last));
    s.f[(K + D) % (D + 1)] = s.R[4 * rot_row<P>(K + D).off];
  }
  __builtin_amdgcn_sched_barrier(0);
  const float f = s.f[K % (D + 1)];
  if constexpr (rw.stage != 1) {
    constexpr int n = rw.blk;
    if constexpr (rw.stage == 0 && rw.kind == 0 && rw.a == 0 && n > 0) rot_z<P, false>(R4 + rot_after(s.last), v, n);
    const float acc = rot_ydot<n, rw.kind>(f, v, std::make_integer_sequence<int, rw.len>{});
    s.tmp[rw.a] = acc;
    s.last = acc;
    if constexpr (rw.a == rw.len - 1) {
#pragma unroll
      for (int a = 0; a < rw.len; ++a) v[rot_cs(n, rw.kind, a)] = s.tmp[a];
      if constexpr (rw.stage == 2 && n > 0 && rw.kind == 1) rot_z<P, true>(R4 + rot_after(acc), v, n);
    }
  } else {
    constexpr int m = rw.blk;
    float sr, si = 0.f;
    rot_adot<m>(f, v, sr, si, std::make_integer_sequence<int, rw.len>{});
    s.tmp[rw.a] = sr;
    s.ti[rw.a] = si;
    s.last = sr;
    if constexpr (rw.a == rw.len - 1) {
#pragma unroll
      for (int a = 0; a < rw.len; ++a) {
        v[rot_re(m + a, m)] = s.tmp[a];
        if (m > 0) v[rot_im(m + a, m)] = s.ti[a];
      }
    }
  }
  __builtin_amdgcn_sched_barrier(0);
}

template <int P, int... J>
__device__ __forceinline__ void rot_prologue(RotPipe<P>& s, std::integer_sequence<int, J...>) {
  ((s.f[J] = s.R[4 * rot_row<P>(J).off]), ...);
}

template <int P, int... K>
__device__ __forceinline__ void rot_run(const float4* __restrict__ R4, float* v, std::integer_sequence<int, K...>) {
  RotPipe<P> s;
  s.R = (rot_lds_float*)reinterpret_cast<const float*>(R4) + (threadIdx.x & 15);
  rot_prologue<P>(s, std::make_integer_sequence<int, RotPipe<P>::D>{});
  (rot_step<P, K>(R4, v, s), ...);
}

// v <- H^T T(|o| z) D v (D = Y_D Z, H^T = Z^T Y_H^T) with the factors of one offset at R4 (LDS, followed by
// at least 16 floats of padding: each lane reads up to 15 floats past a row's start)
template <int P>
__device__ __forceinline__ void rot_apply(const float4* __restrict__ R4, float* v) {
  rot_run<P>(R4, v, std::make_integer_sequence<int, rot_rows<P>()>{});
}

// P2M matrices, built on the GPU (plan.cpp builds the same on the host for the CPU emulator):
// A[layer][col][row] of leaf l = sum over the leaf's faces f and the q Dunavant points (b, w) of f of
// W b_k (layer 0: conj R_n^m(x~), layer 1: conj grad R_n^m(x~) . n / s) for the local column of the face's
// k-th DOF, x~ = (x - centre)/s, W = J w, packed by rows as the P2M kernel reads them (moments.hpp:308-368).
// One workgroup per leaf; per chunk of kP2MPts points, one thread per point builds the unnormalised table
// NR_n^m (n <= P, m >= 0) in LDS by the recurrences NR_m^m = (x+iy)/(2m) NR_{m-1}^{m-1},
// NR_n^m = ((2n-1) z NR_{n-1}^m - r^2 NR_{n-2}^m)/((n+m)(n-m)); then thread (layer, row) adds the chunk's
// contributions to its own row (no two threads write one entry). R_n^m = cnm[n,m] NR_n^m,
// grad R_n^m . n = cnm (-a NR_{n-1}^{m+1} + b NR_{n-1}^{m-1} + n_z NR_{n-1}^m), a = (n_x - i n_y)/2,
// b = (n_x + i n_y)/2, NR_n^{-m} = (-1)^m conj NR_n^m.
constexpr int kP2MPts = 16;
template <int P>
__global__ __launch_bounds__(256) void k_p2m_build(int q, int CP, const int64_t* __restrict__ face_ptr,
                                                   const int32_t* __restrict__ faces, const int32_t* __restrict__ face_lc,
                                                   const double* __restrict__ leaf_cs, const int64_t* __restrict__ mat_ptr,
                                                   const int64_t* __restrict__ col_ptr, const double* __restrict__ fX,
                                                   const double* __restrict__ fcr, const double* __restrict__ fJ,
                                                   const double* __restrict__ rule, const double* __restrict__ cnm,
                                                   double* __restrict__ mat) {
  constexpr int C = (P + 1) * (P + 1);
  constexpr int NT = (P + 1) * (P + 2) / 2;
  __shared__ double nre[kP2MPts][NT], nim[kP2MPts][NT];
  __shared__ double pw[kP2MPts][3], pn[kP2MPts][3];
  __shared__ int plc[kP2MPts][3];
  const int l = blockIdx.x;
  const int ncol = (int)(col_ptr[l + 1] - col_ptr[l]);
  double* M = mat + mat_ptr[l];
  const double cx = leaf_cs[4 * l], cy = leaf_cs[4 * l + 1], cz = leaf_cs[4 * l + 2], s = leaf_cs[4 * l + 3];
  const int tid = threadIdx.x;
  // thread (layer, row) owns rows `row` of its layer in every column: it zeroes them and adds to them
  const int layer = tid >= C ? 1 : 0, row = tid - layer * C;
  const bool owner = tid < 2 * C;
  int rn = 0, rm = 0, rim = 0;  // row -> (n, m, imaginary part)
  if (owner) {
    while ((rn + 1) * (rn + 1) <= row) ++rn;
    const int k = row - rn * rn;
    rm = (k + 1) / 2;
    rim = (k > 0 && (k & 1) == 0) ? 1 : 0;
    for (int c = 0; c < ncol; ++c) M[(int64_t)(layer * ncol + c) * CP + row] = 0.;
  }
  auto tn = [](int n, int m) { return n * (n + 1) / 2 + m; };
  const int64_t f0 = face_ptr[l];
  const int npts = (int)(face_ptr[l + 1] - f0) * q;
  for (int p0 = 0; p0 < npts; p0 += kP2MPts) {
    const int np = min(kP2MPts, npts - p0);
    if (tid < np) {
      const int fi = (p0 + tid) / q, j = (p0 + tid) - fi * q;
      const int f = faces[f0 + fi];
      const double b0 = rule[4 * j], b1 = rule[4 * j + 1], b2 = rule[4 * j + 2], w = rule[4 * j + 3];
      const double J = fJ[f];
      const double x = (b0 * fX[9 * f] + b1 * fX[9 * f + 3] + b2 * fX[9 * f + 6] - cx) / s;
      const double y = (b0 * fX[9 * f + 1] + b1 * fX[9 * f + 4] + b2 * fX[9 * f + 7] - cy) / s;
      const double z = (b0 * fX[9 * f + 2] + b1 * fX[9 * f + 5] + b2 * fX[9 * f + 8] - cz) / s;
      const double W = J * w;
      pw[tid][0] = W * b0;
      pw[tid][1] = W * b1;
      pw[tid][2] = W * b2;
      for (int a = 0; a < 3; ++a) {
        plc[tid][a] = face_lc[3 * (f0 + fi) + a];
        pn[tid][a] = fcr[3 * f + a] / J;
      }
      double* re = nre[tid];
      double* im = nim[tid];
      const double r2 = x * x + y * y + z * z;
      re[0] = 1.;
      im[0] = 0.;
      for (int m = 1; m <= P; ++m) {
        const double pr = re[tn(m - 1, m - 1)], pi = im[tn(m - 1, m - 1)], sc = 1. / (2. * m);
        re[tn(m, m)] = sc * (x * pr - y * pi);
        im[tn(m, m)] = sc * (x * pi + y * pr);
      }
      for (int m = 0; m <= P; ++m)
        for (int n = m + 1; n <= P; ++n) {
          double ar = (2 * n - 1) * z * re[tn(n - 1, m)], ai = (2 * n - 1) * z * im[tn(n - 1, m)];
          if (n - 2 >= m) {
            ar -= r2 * re[tn(n - 2, m)];
            ai -= r2 * im[tn(n - 2, m)];
          }
          const double sc = 1. / ((double)(n + m) * (n - m));
          re[tn(n, m)] = sc * ar;
          im[tn(n, m)] = sc * ai;
        }
    }
    __syncthreads();
    if (owner) {
      const double c = cnm[tn(rn, rm)];
      for (int pt = 0; pt < np; ++pt) {
        double v;
        if (layer == 0) {
          v = rim ? -c * nim[pt][tn(rn, rm)] : c * nre[pt][tn(rn, rm)];
        } else if (rn == 0) {
          v = 0.;
        } else {
          // NR_{n-1}^{q}: zero if |q| > n-1; negative q from the conjugate
          auto nr = [&](int n, int qq, double& r, double& i) {
            if (n < 0 || qq > n || -qq > n) {
              r = i = 0.;
              return;
            }
            if (qq >= 0) {
              r = nre[pt][tn(n, qq)];
              i = nim[pt][tn(n, qq)];
            } else {
              const double sg = ((-qq) & 1) ? -1. : 1.;
              r = sg * nre[pt][tn(n, -qq)];
              i = -sg * nim[pt][tn(n, -qq)];
            }
          };
          double p1r, p1i, m1r, m1i, z0r, z0i;
          nr(rn - 1, rm + 1, p1r, p1i);
          nr(rn - 1, rm - 1, m1r, m1i);
          nr(rn - 1, rm, z0r, z0i);
          const double nx = pn[pt][0], ny = pn[pt][1], nz = pn[pt][2];
          const double ar = 0.5 * nx, ai = -0.5 * ny, br = 0.5 * nx, bi = 0.5 * ny;
          const double dr = -(ar * p1r - ai * p1i) + (br * m1r - bi * m1i) + nz * z0r;
          const double di = -(ar * p1i + ai * p1r) + (br * m1i + bi * m1r) + nz * z0i;
          v = rim ? -(c / s) * di : (c / s) * dr;
        }
        for (int k = 0; k < 3; ++k) M[(int64_t)(layer * ncol + plc[pt][k]) * CP + row] += pw[pt][k] * v;
      }
    }
    __syncthreads();
  }
}

// FP32 multipoles for the rotation M2L: M32[(box*2 + layer)*C + r] = M_layer[box*C + r]
__global__ void k_pack_m32(int64_t n, int C, const double* __restrict__ M0, const double* __restrict__ M1,
                           float* __restrict__ M32) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t col = i / C, r = i - col * C;
    M32[i] = (float)((col & 1) ? M1 : M0)[(col >> 1) * C + r];
  }
}

template <int P>
__global__ __launch_bounds__(256) void k_m2l_rot(int ntiles, const float* __restrict__ rot,
                                                 const int32_t* __restrict__ tile_op, const int64_t* __restrict__ tile_p0,
                                                 const int64_t* __restrict__ tile_p1, const int32_t* __restrict__ src,
                                                 const float* __restrict__ M32, float* __restrict__ stage) {
  constexpr int C = (P + 1) * (P + 1);
  constexpr int RS = rot_size(P);
  constexpr int NCH = (C + 15) / 16;
  constexpr int kS = 17;  // LDS column stride (odd: conflict-free column reads)
  __shared__ float4 R4[RS / 4 + 4];  // + 16 floats read past the last row by lanes 4q..15
  __shared__ float Tw[4][64 * kS];
  const int per_xcd = (ntiles + kXCD - 1) / kXCD;
  const int t = (blockIdx.x % kXCD) * per_xcd + blockIdx.x / kXCD;
  if (t >= ntiles) return;
  const int op = tile_op[t];
  const int64_t p0 = tile_p0[t];
  const int64_t p1 = tile_p1[t];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t pw = p0 + wave * kRotWavePairs;
  const int np = (int)max<int64_t>(0, min<int64_t>(kRotWavePairs, p1 - pw));
  // ---- issue the gather loads first: lane (g, k) loads rows 16 ch + k of columns 4i + g ----
  const int g = lane >> 4, k = lane & 15;
  float ld[NCH][16];
  if (np > 0) {
    const int my_src = src[pw + min(lane >> 1, np - 1)];  // columns past the tile repeat its last pair
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int col = 4 * i + g;
      const float* S = M32 + ((int64_t)__shfl(my_src, col) * 2 + (col & 1)) * C;
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch) ld[ch][i] = (ch * 16 + k < C) ? S[ch * 16 + k] : 0.f;
    }
  }
  // ---- the offset's factors -> LDS (shared by the 4 waves) ----
  const float4* g4 = reinterpret_cast<const float4*>(rot + (int64_t)op * RS);
  for (int i = threadIdx.x; i < RS / 4; i += 256) R4[i] = g4[i];  // RS = 4 * rot_size4(P)
  __syncthreads();
  if (np <= 0) return;
  float* T = Tw[wave];
  float v[C];
  // ---- transpose the gathered rows into the lane's column, 16 rows at a time ----
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
#pragma unroll
    for (int i = 0; i < 16; ++i) T[(4 * i + g) * kS + k] = ld[ch][i];
    wave_lds_sync();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk)
      if (ch * 16 + kk < C) v[ch * 16 + kk] = T[lane * kS + kk];
    wave_lds_sync();
  }
  // ---- D, axial M2L, H^T ----
  __builtin_amdgcn_sched_barrier(0);
  rot_apply<P>(R4, v);
  __builtin_amdgcn_sched_barrier(0);
  // ---- scatter: lane's registers -> LDS -> rows [16 ch, 16 ch + 16) of column (4i + g) ----
  constexpr int CS = stage_stride<float>(C);
  static_assert(16 * NCH == CS, "the staged columns are written 16 rows at a time");
  float* out = stage + pw * 2 * CS;
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) T[lane * kS + kk] = ch * 16 + kk < C ? v[ch * 16 + kk] : 0.f;
    wave_lds_sync();
    const int r = ch * 16 + k;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int col = 4 * i + g;
      if (col < 2 * np) out[(int64_t)col * CS + r] = T[col * kS + k];
    }
    wave_lds_sync();
  }
}

// FP32 staging: the same reduction with 16-byte loads. Thread t < CS/2 of wave w reads float4 t of a
// pair's two staged columns (rows 4t..4t+3 of one layer, CS = stage_stride<float>(C)); wave w takes the
// pairs a + w, a + w + 4, ... (4 in flight per thread) and the four waves' sums are combined in a fixed
// order (deterministic). Boxes run in descending pair count (order), as in k_m2l_reduce.
__global__ __launch_bounds__(256) void k_m2l_reduce_f32(int C, const int32_t* __restrict__ order,
                                                        const int64_t* __restrict__ tptr,
                                                        const int32_t* __restrict__ tpairs,
                                                        const float* __restrict__ stage, double* __restrict__ L0,
                                                        double* __restrict__ L1) {
  __shared__ double part[3][4][64];
  const int CS = stage_stride<float>(C);  // <= 128, a multiple of 32
  const int tb = order[blockIdx.x];
  const int w = threadIdx.x >> 6, t = threadIdx.x & 63;
  const bool act = 4 * t < 2 * CS;
  const int64_t a = tptr[tb], b = tptr[tb + 1];
  double s0 = 0., s1 = 0., s2 = 0., s3 = 0.;
  if (act) {
    int64_t q = a + w;
    for (; q + 12 < b; q += 16) {  // 4 loads in flight
      float4 v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = ld_nt(reinterpret_cast<const float4*>(stage + (int64_t)tpairs[q + 4 * u] * 2 * CS) + t);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        s0 += v[u].x;
        s1 += v[u].y;
        s2 += v[u].z;
        s3 += v[u].w;
      }
    }
    for (; q < b; q += 4) {
      const float4 v = ld_nt(reinterpret_cast<const float4*>(stage + (int64_t)tpairs[q] * 2 * CS) + t);
      s0 += v.x;
      s1 += v.y;
      s2 += v.z;
      s3 += v.w;
    }
  }
  if (w > 0 && act) {
    part[w - 1][0][t] = s0;
    part[w - 1][1][t] = s1;
    part[w - 1][2][t] = s2;
    part[w - 1][3][t] = s3;
  }
  __syncthreads();
  if (w == 0 && act) {
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      s0 += part[v][0][t];
      s1 += part[v][1][t];
      s2 += part[v][2][t];
      s3 += part[v][3][t];
    }
    const int f = 4 * t;  // a float4 never straddles the two layers (CS is a multiple of 4)
    const int layer = f >= CS ? 1 : 0, r = f - layer * CS;
    double* L = (layer ? L1 : L0) + (int64_t)tb * C;
    if (r < C) L[r] = s0;
    if (r + 1 < C) L[r + 1] = s1;
    if (r + 2 < C) L[r + 2] = s2;
    if (r + 3 < C) L[r + 3] = s3;
  }
}

// FP64 -> FP32 copy of a resident matrix (FP32 near field / P2M / operators).
__global__ void k_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) b[i] = (float)a[i];
}

// Av_i = GPhin - (GnPhi + alpha_i phi_i) (CORNER Neumann rows: Av_i = phi_i). mode 1: b = -Av (CORNER: phi_D).
__global__ void k_assemble(int mode, int n, const int32_t* __restrict__ kind, const int32_t* __restrict__ point,
                           const double* __restrict__ alpha, const double* __restrict__ phi, const double* __restrict__ y0,
                           const double* __restrict__ y1, const double* __restrict__ phi_k, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int p = point[i];
  if (mode == 2) {
    double a = -y1[p];
    if (fabs(a) < 1e-15) a = 1.0;
    out[i] = a;
    return;
  }
  double v = (kind[i] == 2) ? phi[i] : y0[p] - (y1[p] + alpha[i] * phi[i]);
  if (mode == 1) v = (kind[i] == 2) ? phi_k[i] : -v;
  out[i] = v;
}

// Near-field quadrature: one workgroup per near row (target node); threads stride over the
// faces of the target box's near ranges and scatter (WG, WGn) of the three corners into the row.
// fill_direct_linear / fill_nonadj_linear (BEM_solveBVP_GMRES_FMM.hpp:1824-1891).
__global__ __launch_bounds__(256) void k_near_quad(
    int C, int q, double near_region, const int32_t* __restrict__ row_block, const int64_t* __restrict__ row_ptr,
    const int64_t* __restrict__ col_ptr, const int64_t* __restrict__ mat_ptr, const int32_t* __restrict__ row_out,
    const int32_t* __restrict__ col_idx, const int32_t* __restrict__ ncol_dof, const int64_t* __restrict__ range_ptr,
    const int64_t* __restrict__ ranges, const int64_t* __restrict__ range_prefix, const int32_t* __restrict__ face_perm,
    const double* __restrict__ Xs, const double* __restrict__ fX, const double* __restrict__ fcent,
    const double* __restrict__ fcr, const double* __restrict__ fJ, const int32_t* __restrict__ fv,
    const int32_t* __restrict__ fdof, const double* __restrict__ l2p, const double* __restrict__ dun_rule,
    double* __restrict__ mat) {
  // Dunavant rule of the non-adjacent, non-near faces ({b0,b1,b2,w} x q), per context
  __shared__ double dun[7][4];
  for (int i = threadIdx.x; i < 28; i += blockDim.x) dun[i / 4][i % 4] = dun_rule[i];
  __syncthreads();
  const int64_t r = blockIdx.x;
  const int b = row_block[r];
  const int ncd = ncol_dof[b];
  const int64_t c0 = col_ptr[b];
  const int ncol = (int)(col_ptr[b + 1] - c0);
  double* A0 = mat + mat_ptr[b] + (r - row_ptr[b]) * 2 * ncol;
  double* A1 = A0 + ncol;
  const int t = row_out[r];
  for (int c = threadIdx.x; c < C; c += blockDim.x) {
    A0[ncd + c] = l2p[r * C + c];
    A1[ncd + c] = l2p[r * C + c];
  }
  const double x = Xs[3 * t], y = Xs[3 * t + 1], z = Xs[3 * t + 2];
  const int64_t g0 = range_ptr[b], g1 = range_ptr[b + 1];
  const int64_t nf = range_prefix[g1 - 1] + (ranges[2 * (g1 - 1) + 1] - ranges[2 * (g1 - 1)]);
  const int32_t* cols = col_idx + c0;
  for (int64_t i = threadIdx.x; i < nf; i += blockDim.x) {
    // locate the range holding local face i
    int64_t lo = g0, hi = g1 - 1;
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (range_prefix[mid] <= i) lo = mid;
      else hi = mid - 1;
    }
    const int f = face_perm[ranges[2 * lo] + (i - range_prefix[lo])];
    int slot = -1;
    for (int k = 0; k < 3; ++k)
      if (fv[3 * f + k] == t) slot = k;
    double WG[3], WGn[3];
    bemfmm::face_weights(fX + 9 * (int64_t)f, fJ[f], fcr + 3 * (int64_t)f, fcent + 3 * (int64_t)f, slot, x, y, z,
                         near_region, c_gw5, dun, q, WG, WGn);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int d = fdof[3 * f + k];
      int lo2 = 0, hi2 = ncd - 1;
      while (lo2 < hi2) {
        const int mid = (lo2 + hi2) >> 1;
        if (cols[mid] < d) lo2 = mid + 1;
        else hi2 = mid;
      }
      atomicAdd(&A0[lo2], WG[k]);
      atomicAdd(&A1[lo2], WGn[k]);
    }
  }
}

// Diagonal of the BIE operator per DOF, read from the near-field row of its point:
// Dirichlet (unknown phi_n): G_ii; Neumann (unknown phi): -Gn_ii - alpha_i; ROBIN: kappa_i G_ii - Gn_ii -
// alpha_i; CORNER rows: 1. dinv = 1/diag (1 where the diagonal vanishes).
__global__ void k_diag_inv(int n, const int32_t* __restrict__ kind, const double* __restrict__ kappa,
                           const int32_t* __restrict__ point,
                           const int64_t* __restrict__ node_row, const int32_t* __restrict__ row_block,
                           const int64_t* __restrict__ row_ptr, const int64_t* __restrict__ col_ptr,
                           const int64_t* __restrict__ mat_ptr, const int32_t* __restrict__ col_idx,
                           const int32_t* __restrict__ ncol_dof, const double* __restrict__ mat,
                           const double* __restrict__ alpha, double* __restrict__ dinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double dg = 1.0;
  if (kind[i] != 2) {
    const int64_t r = node_row[point[i]];
    const int b = row_block[r];
    const int64_t c0 = col_ptr[b];
    const int ncol = (int)(col_ptr[b + 1] - c0);
    const int ncd = ncol_dof[b];
    int lo = 0, hi = ncd - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (col_idx[c0 + mid] < i) lo = mid + 1;
      else hi = mid;
    }
    const double* A0 = mat + mat_ptr[b] + (r - row_ptr[b]) * 2 * ncol;
    dg = (kind[i] == 0) ? A0[lo] : -A0[ncol + lo] - alpha[i];
    if (kind[i] == 3) dg += kappa[i] * A0[lo];
  }
  dinv[i] = (isfinite(dg) && fabs(dg) > 1e-15) ? 1.0 / dg : 1.0;
}

// ---- restricted additive Schwarz preconditioner ----
// Operator entry (row DOF i, column DOF j) from the resident near field: the near row of i's point,
// column j if it is in that row's near set (else 0); G for Dirichlet columns, kappa_j G - Gn for ROBIN
// columns, -Gn otherwise; -alpha_i on Neumann and ROBIN diagonals; CORNER rows are e_i.
// (buildSparseMatrixForILU, GMRES_FMM.hpp:440-640)
__device__ double near_entry(int i, int j, const int32_t* __restrict__ kind, const double* __restrict__ kappa,
                             const int32_t* __restrict__ point,
                             const int64_t* __restrict__ node_row, const int32_t* __restrict__ row_block,
                             const int64_t* __restrict__ row_ptr, const int64_t* __restrict__ col_ptr,
                             const int64_t* __restrict__ mat_ptr, const int32_t* __restrict__ col_idx,
                             const int32_t* __restrict__ ncol_dof, const double* __restrict__ mat,
                             const double* __restrict__ alpha) {
  const int ki = kind[i];
  if (ki == 2) return i == j ? 1.0 : 0.0;
  const int64_t r = node_row[point[i]];
  const int b = row_block[r];
  const int64_t c0 = col_ptr[b];
  const int ncol = (int)(col_ptr[b + 1] - c0);
  const int ncd = ncol_dof[b];
  int lo = 0, hi = ncd - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (col_idx[c0 + mid] < j) lo = mid + 1;
    else hi = mid;
  }
  double v = 0.;
  if (ncd > 0 && col_idx[c0 + lo] == j) {
    const double* A0 = mat + mat_ptr[b] + (r - row_ptr[b]) * 2 * ncol;
    v = (kind[j] == 0) ? A0[lo] : -A0[ncol + lo];
    if (kind[j] == 3) v += kappa[j] * A0[lo];
  }
  if (i == j && (ki == 1 || ki == 3)) v -= alpha[i];
  return v;
}

// One workgroup per block (m <= 128 DOFs): extract A_b into LDS, invert it in place by Gauss-Jordan
// elimination with partial pivoting, and store the core rows of A_b^-1 (ncore x m, row-major).
// A block with a vanishing or non-finite pivot falls back to the inverse diagonal on its core
// (as SchwarzAdditive_CRS does when its block LU fails) and is flagged in failed[b].
// LDS: m*m doubles (matrix) + m doubles (pivot column) + m ints (DOF ids) + m ints (pivot rows).
constexpr int kSchwarzMaxBlock = 128;
static_assert(kSchwarzMaxBlock <= 128, "k_schwarz_build updates one column per thread, two rows per pass of 256 threads");
__global__ __launch_bounds__(256) void k_schwarz_build(
    const int64_t* __restrict__ bptr, const int32_t* __restrict__ bidx, const int32_t* __restrict__ bncore,
    const int64_t* __restrict__ inv_ptr, const int32_t* __restrict__ kind, const double* __restrict__ kappa,
    const int32_t* __restrict__ point, const int64_t* __restrict__ node_row, const int32_t* __restrict__ row_block,
    const int64_t* __restrict__ row_ptr,
    const int64_t* __restrict__ col_ptr, const int64_t* __restrict__ mat_ptr, const int32_t* __restrict__ col_idx,
    const int32_t* __restrict__ ncol_dof, const double* __restrict__ mat, const double* __restrict__ alpha,
    double* __restrict__ inv, int32_t* __restrict__ failed) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int b = blockIdx.x;
  const int64_t p0 = bptr[b];
  const int m = (int)(bptr[b + 1] - p0);
  const int nc = bncore[b];
  double* A = reinterpret_cast<double*>(smem);   // m*m
  double* colk = A + m * m;                      // m
  int* ids = reinterpret_cast<int*>(colk + m);   // m
  int* piv = ids + m;                            // m
  __shared__ int bad;
  const int tid = threadIdx.x;
  if (tid == 0) bad = 0;
  for (int t = tid; t < m; t += blockDim.x) ids[t] = bidx[p0 + t];
  __syncthreads();
  for (int e = tid; e < m * m; e += blockDim.x)
    A[e] = near_entry(ids[e / m], ids[e % m], kind, kappa, point, node_row, row_block, row_ptr, col_ptr, mat_ptr,
                      col_idx, ncol_dof, mat, alpha);
  __syncthreads();
  for (int k = 0; k < m; ++k) {
    if (tid < 64) {
      // pivot: largest |A[i][k]| over i >= k (lowest row on ties)
      double best = -1.0;
      int bi = k;
      for (int i = k + tid; i < m; i += 64) {
        const double v = fabs(A[i * m + k]);
        if (v > best) {
          best = v;
          bi = i;
        }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double ob = __shfl_down(best, off);
        const int oi = __shfl_down(bi, off);
        if (ob > best || (ob == best && oi < bi)) {
          best = ob;
          bi = oi;
        }
      }
      if (tid == 0) piv[k] = bi;
    }
    __syncthreads();
    const int pr = piv[k];
    if (pr != k)
      for (int j = tid; j < m; j += blockDim.x) {
        const double t = A[k * m + j];
        A[k * m + j] = A[pr * m + j];
        A[pr * m + j] = t;
      }
    __syncthreads();
    const double pv = A[k * m + k];
    if (!(fabs(pv) > 1e-300) || !isfinite(pv)) {
      if (tid == 0) bad = 1;
      break;  // uniform: every thread reads the same pivot
    }
    const double ipv = 1.0 / pv;
    for (int i = tid; i < m; i += blockDim.x) colk[i] = A[i * m + k];
    __syncthreads();
    for (int j = tid; j < m; j += blockDim.x) A[k * m + j] = (j == k) ? ipv : A[k * m + j] * ipv;
    __syncthreads();
    {  // one column per thread, two rows per pass (m <= 128 = blockDim.x / 2): consecutive threads take
       // consecutive columns of a row (no bank conflicts) and no thread divides by m
      const int j = tid & 127;
      if (j < m) {
        const double akj = A[k * m + j];  // the scaled pivot row (row k is not updated here)
        for (int i = tid >> 7; i < m; i += 2) {
          if (i == k) continue;
          A[i * m + j] = (j == k) ? -colk[i] * ipv : A[i * m + j] - colk[i] * akj;
        }
      }
    }
    __syncthreads();
  }
  __syncthreads();
  double* out = inv + inv_ptr[b];
  if (bad) {
    if (tid == 0) failed[b] = 1;
    for (int e = tid; e < nc * m; e += blockDim.x) {
      const int i = e / m, j = e % m;
      double v = 0.;
      if (i == j) {
        const double d = near_entry(ids[i], ids[i], kind, kappa, point, node_row, row_block, row_ptr, col_ptr,
                                    mat_ptr, col_idx, ncol_dof, mat, alpha);
        v = (isfinite(d) && fabs(d) > 1e-15) ? 1.0 / d : 1.0;
      }
      out[e] = v;
    }
    return;
  }
  if (tid == 0) failed[b] = 0;
  // undo the row interchanges: swap columns in reverse order
  for (int k = m - 1; k >= 0; --k) {
    const int pr = piv[k];
    if (pr != k)
      for (int i = tid; i < m; i += blockDim.x) {
        const double t = A[i * m + k];
        A[i * m + k] = A[i * m + pr];
        A[i * m + pr] = t;
      }
    __syncthreads();
  }
  for (int e = tid; e < nc * m; e += blockDim.x) out[e] = A[e];
}

// z[core DOFs of b] = (A_b^-1 r|_b)[core]: one workgroup per block; each DOF is written by exactly
// the one block whose core holds it (restricted additive Schwarz), so z needs no clearing or atomics.
__global__ __launch_bounds__(256) void k_schwarz_apply(const int64_t* __restrict__ bptr, const int32_t* __restrict__ bidx,
                                                       const int32_t* __restrict__ bncore, const int64_t* __restrict__ inv_ptr,
                                                       const double* __restrict__ inv, const double* __restrict__ r,
                                                       double* __restrict__ z) {
  __shared__ double rl[kSchwarzMaxBlock];
  const int b = blockIdx.x;
  const int64_t p0 = bptr[b];
  const int m = (int)(bptr[b + 1] - p0);
  const int nc = bncore[b];
  for (int t = threadIdx.x; t < m; t += blockDim.x) rl[t] = r[bidx[p0 + t]];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const double* Ib = inv + inv_ptr[b];
  for (int i = wave; i < nc; i += 4) {
    const double* row = Ib + (int64_t)i * m;
    double s = 0.;
    for (int j = lane; j < m; j += 64) s += row[j] * rl[j];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
    if (lane == 0) z[bidx[p0 + i]] = s;
  }
}

// ---- Morton DOF order (GMRES_FMM.hpp:3164-3246) ----
// 21 bits per axis, interleaved x | y<<1 | z<<2: spread the 21 low bits of v to every third bit
__device__ __forceinline__ uint64_t spread3(uint32_t v) {
  uint64_t x = v & 0x1fffffu;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}
// key of each DOF from its point, quantised in the bounding box of all points as the reference does
// (t = (v - v0)/(v1 - v0) clamped to [0, 1], q = (uint32)(t * (2^21 - 1) + 0.5)); vals = DOF index
__global__ void k_morton_keys(int n, const int32_t* __restrict__ point, const double* __restrict__ X, double x0, double x1,
                             double y0, double y1, double z0, double z1, uint64_t* __restrict__ keys,
                             int32_t* __restrict__ vals) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int p = point[i];
  auto q = [](double v, double a, double b) -> uint32_t {
    if (!(b > a)) return 0u;
    double t = (v - a) / (b - a);
    t = fmin(fmax(t, 0.0), 1.0);
    return (uint32_t)(t * (double)((1u << 21) - 1u) + 0.5);
  };
  keys[i] = spread3(q(X[3 * p], x0, x1)) | (spread3(q(X[3 * p + 1], y0, y1)) << 1) |
            (spread3(q(X[3 * p + 2], z0, z1)) << 2);
  vals[i] = i;
}
// caller order <-> internal (Morton) order: internal i holds caller DOF perm[i]
__global__ void k_perm_gather(int n, const int32_t* __restrict__ perm, const double* __restrict__ in,
                              double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[perm[i]];
}
__global__ void k_perm_scatter(int n, const int32_t* __restrict__ perm, const double* __restrict__ in,
                               double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[perm[i]] = in[i];
}

// ---- GMRES vector kernels (FP64) ----
constexpr int kDotBlocks = 256;
// partial[i * kDotBlocks + blk] = sum over the block's chunk of V_i . w   (grid: kDotBlocks x k)
__global__ __launch_bounds__(256) void k_dots(const double* __restrict__ V, int64_t ldv, const double* __restrict__ w, int64_t n,
                                              double* __restrict__ partial) {
  __shared__ double red[4];
  const int i = blockIdx.y;
  const double* v = V + i * ldv;
  double s = 0.;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) s += v[j] * w[j];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partial[(int64_t)i * gridDim.x + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}
__global__ void k_reduce_partials(const double* __restrict__ partial, int nblk, int k, double* __restrict__ out) {
  const int i = blockIdx.x;
  __shared__ double red[4];
  double s = 0.;
  for (int j = threadIdx.x; j < nblk; j += blockDim.x) s += partial[(int64_t)i * nblk + j];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[i] = red[0] + red[1] + red[2] + red[3];
  (void)k;
}
// w[j] += sign * sum_{i<k} V_i[j] h[i]
__global__ void k_vsum(const double* __restrict__ V, int64_t ldv, int k, const double* __restrict__ h, double sign,
                       double* __restrict__ w, int64_t n) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.;
    for (int i = 0; i < k; ++i) s += V[i * ldv + j] * h[i];
    w[j] += sign * s;
  }
}
__global__ void k_axpby(int64_t n, double a, const double* __restrict__ x, double b, double* __restrict__ y) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) y[j] = a * x[j] + b * y[j];
}
__global__ void k_sub(int64_t n, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ out) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) out[j] = a[j] - b[j];
}
__global__ void k_mul(int64_t n, const double* __restrict__ d, double* __restrict__ y) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) y[j] *= d[j];
}
__global__ void k_mul_to(int64_t n, const double* __restrict__ d, const double* __restrict__ x, double* __restrict__ y) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) y[j] = d[j] * x[j];
}
// w *= a, the result also written to dup
__global__ void k_scale_dup(int64_t n, double a, double* __restrict__ w, double* __restrict__ dup) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    const double v = a * w[j];
    w[j] = v;
    dup[j] = v;
  }
}

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------
template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    release();
    if (count == 0) count = 1;
    HIPCHECK(hipMalloc((void**)&p, count * sizeof(T)));
    n = count;
  }
  // Setup-time upload; completes before returning (the source may be a temporary).
  void upload(const std::vector<T>& v, hipStream_t s) {
    alloc(v.size());
    if (!v.empty()) {
      HIPCHECK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
      HIPCHECK(hipStreamSynchronize(s));
    }
  }
  size_t bytes() const { return n * sizeof(T); }
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
};

int grid1(int64_t n, int bs = 256) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + bs - 1) / bs, 65536)); }

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
  DBuf<float> near_mat32, p2m_mat32, ops32;
  // device operator table, [slot][KP][RP], kept across prepares: slot of each (kind, a, b, c) key
  std::map<std::array<int, 4>, int32_t> op_slot;
  int ops_cap = 0, ops_used = 0, ops_P = -1, ops32_valid = 0;
  int ops_uploaded = 0;  // operators uploaded by the last prepare
  std::vector<int32_t> m2l_slot;  // plan op -> table slot (for re-tiling by the autotuner)
  int m2l_tp = 0;                // pairs per M2L tile in use
  // rotation M2L of the FP32 operator: factors [slot][RS] (slots of the operator table; M2L slots only)
  bool m2l_rot = false;
  DBuf<float> rot32;
  std::vector<char> rot_have;
  int rot_RS = 0;
  DTiles m2l_rt;  // M2L tiles of kRotTP pairs (one offset each)
  int m2l_min_level = 0;  // coarsest level with M2L pairs: M2M into and L2L out of coarser levels are skipped
  DBuf<float> m32;  // FP32 multipoles [box][layer][C] read by the rotation M2L
  // near field beside the FMM passes: the near-only GEMV runs on stream2 between a fork and a join event
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
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
  int sw_failed = 0;
  DBuf<int64_t> sw_ptr, sw_invptr;
  DBuf<int32_t> sw_idx, sw_ncore;
  DBuf<double> sw_inv;
  DBuf<double> V;            // Krylov basis
  int64_t v_cap = 0;
  DBuf<double> partial, hbuf;
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

template <typename T>
void launch_gemv(bemfmm_ctx* c, DGemv& g, const T* mat, const double* x0, const double* x1, double* y0, double* y1,
                 int part = 0, hipStream_t s = nullptr) {
  if (g.nblk == 0) return;
  hipLaunchKernelGGL(k_block_gemv<T>, dim3(g.nblk), dim3(256), 0, s ? s : c->stream, g.row_ptr.p, g.col_ptr.p,
                     g.mat_ptr.p, g.row_out.p, g.col_idx.p, mat, x0, x1, y0, y1, part, c->near_ncd.p);
}

template <typename T>
void launch_translate(bemfmm_ctx* c, DTiles& t, const T* ops, const double* S0, const double* S1, double* D0, double* D1,
                      T* stage = nullptr) {
  if (t.ntiles == 0) return;
  const size_t lds = translate_lds(t.tp, c->KP, c->RP) * sizeof(T);
  const int grid = ((t.ntiles + kXCD - 1) / kXCD) * kXCD;
  const int RTW = c->RP / 64;
#define BEMFMM_TR_TP(N, TPV)                                                                                            \
  if (stage)                                                                                                            \
    hipLaunchKernelGGL((k_translate_mfma<T, N, true, TPV>), dim3(grid), dim3(256), lds, c->stream, c->C, c->RP, c->KP,  \
                       t.ntiles, ops, t.op.p, t.p0.p, t.p1.p, t.dst.p, t.src.p, S0, S1, D0, D1, stage);                  \
  else                                                                                                                  \
    hipLaunchKernelGGL((k_translate_mfma<T, N, false, TPV>), dim3(grid), dim3(256), lds, c->stream, c->C, c->RP, c->KP, \
                       t.ntiles, ops, t.op.p, t.p0.p, t.p1.p, t.dst.p, t.src.p, S0, S1, D0, D1, stage)
#define BEMFMM_TR(N)                                      \
  do {                                                    \
    if (!stage && t.tp == 8) BEMFMM_TR_TP(N, 8);          \
    else if (t.tp == 16) BEMFMM_TR_TP(N, 16);             \
    else if (stage && t.tp == 64) BEMFMM_TR_TP(N, 64);    \
    else BEMFMM_TR_TP(N, 32);                             \
  } while (0)
  switch (RTW) {
  case 1: BEMFMM_TR(1); break;
  case 2: BEMFMM_TR(2); break;
  case 3: BEMFMM_TR(3); break;
  case 4: BEMFMM_TR(4); break;
  case 5: BEMFMM_TR(5); break;
  case 6: BEMFMM_TR(6); break;
  default: BEMFMM_TR(7); break;
  }
#undef BEMFMM_TR
#undef BEMFMM_TR_TP
}

bool rot_compiled(int P) { return P == 8 || P == 10; }

void launch_m2l_rot(bemfmm_ctx* c, float* stage) {
  DTiles& t = c->m2l_rt;
  if (t.ntiles == 0) return;
  const int grid = ((t.ntiles + kXCD - 1) / kXCD) * kXCD;
  if (c->rot_RS != rot_size(c->plan->P)) throw std::runtime_error("rotation M2L: factor table stride mismatch");
  const int64_t nm = (int64_t)c->n_boxes * 2 * c->C;
  hipLaunchKernelGGL(k_pack_m32, dim3(grid1(nm)), dim3(256), 0, c->stream, nm, c->C, c->M0.p, c->M1.p, c->m32.p);
#define BEMFMM_ROT(PV)                                                                                          \
  hipLaunchKernelGGL(k_m2l_rot<PV>, dim3(grid), dim3(256), 0, c->stream, t.ntiles, c->rot32.p, t.op.p, t.p0.p, \
                     t.p1.p, t.src.p, c->m32.p, stage)
  switch (c->plan->P) {
  case 8: BEMFMM_ROT(8); break;
  case 10: BEMFMM_ROT(10); break;
  default: throw std::runtime_error("rotation M2L not compiled for this expansion order");
  }
#undef BEMFMM_ROT
}

void fill(bemfmm_ctx* c, double* p, int64_t n, double v) {
  hipLaunchKernelGGL(k_fill, dim3(grid1(n)), dim3(256), 0, c->stream, p, n, v);
}

// One evaluation of the BIE operator. mode: 0 = A x (x = in), 1 = rhs, 2 = rigid-mode alpha.
// lo: FP32 near field, P2M and translations (FP32 MFMA), FP64 vectors and accumulation across boxes.
// Launches only (no host synchronisation), so that it can be captured into a HIP graph.
void evaluate_launch(bemfmm_ctx* c, int mode, const double* in, double* out, bool lo) {
  Plan& pl = *c->plan;
  const int n = c->n_dof;
  const int64_t nbC = (int64_t)c->n_boxes * c->C;
  double* phin = c->buf0.p;
  double* phi = c->buf1.p;
  double* L0 = c->buf0.p + n;
  double* L1 = c->buf1.p + n;
  auto rec = [&](int i) {
    if (c->prof) HIPCHECK(hipEventRecord(c->ev[i], c->stream));
  };
  rec(0);
  hipLaunchKernelGGL(k_scatter, dim3(grid1(n)), dim3(256), 0, c->stream, mode == 0 ? 0 : (mode == 1 ? 1 : 2), n,
                     c->kind.p, c->robin.p, in, c->phi_k.p, c->phin_k.p, phin, phi);
  // the near-field columns need only the scattered DOF values: they run on stream2 beside the FMM passes
  // (not while profiling, which times the stages one after the other)
  const bool split = c->opts.overlap_near && !c->prof && c->near.nblk > 0;
  if (split) {
    HIPCHECK(hipEventRecord(c->ev_fork, c->stream));
    HIPCHECK(hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
    if (lo) launch_gemv(c, c->near, c->near_mat32.p, c->buf0.p, c->buf1.p, c->y0.p, c->y1.p, 1, c->stream2);
    else launch_gemv(c, c->near, c->near.mat.p, c->buf0.p, c->buf1.p, c->y0.p, c->y1.p, 1, c->stream2);
    HIPCHECK(hipEventRecord(c->ev_join, c->stream2));
  }
  fill(c, c->M0.p, nbC, 0.);
  fill(c, c->M1.p, nbC, 0.);
  rec(1);
  if (c->p2m.nblk > 0) {
    if (lo)
      hipLaunchKernelGGL(k_p2m<float>, dim3(c->p2m.nblk), dim3(256), 0, c->stream, c->C, c->p2m_rows_pad,
                         c->p2m.row_ptr.p, c->p2m.col_ptr.p, c->p2m.mat_ptr.p, c->p2m.row_out.p, c->p2m.col_idx.p,
                         c->p2m_mat32.p, phin, phi, c->M0.p, c->M1.p);
    else
      hipLaunchKernelGGL(k_p2m<double>, dim3(c->p2m.nblk), dim3(256), 0, c->stream, c->C, c->p2m_rows_pad,
                         c->p2m.row_ptr.p, c->p2m.col_ptr.p, c->p2m.mat_ptr.p, c->p2m.row_out.p, c->p2m.col_idx.p,
                         c->p2m.mat.p, phin, phi, c->M0.p, c->M1.p);
  }
  rec(2);
  // M2M from level l into l-1 only where level l-1 has M2L pairs (coarser multipoles are not used).
  // M2M and L2L use the FP64 operators for both operators: their inputs and outputs are FP64, and the
  // FP64 matrix-core tiles ran faster here than the FP32 ones (0.115 / 0.120 ms against 0.157 / 0.161 ms
  // per matvec at 209k nodes)
  for (int l = pl.n_levels - 1; l >= 1 && l - 1 >= c->m2l_min_level; --l) {
    launch_translate(c, c->m2m[l], c->ops.p, c->M0.p, c->M1.p, c->M0.p, c->M1.p);
  }
  rec(3);
  if (lo) {
    float* st = reinterpret_cast<float*>(c->m2l_stage.p);
    if (c->m2l_rot) launch_m2l_rot(c, st);
    else launch_translate(c, c->m2l, c->ops32.p, c->M0.p, c->M1.p, nullptr, nullptr, st);
    rec(4);
    hipLaunchKernelGGL(k_m2l_reduce_f32, dim3(c->n_boxes), dim3(256), 0, c->stream, c->C, c->m2l_rorder.p, c->m2l_tptr.p,
                       c->m2l_tpairs.p, (const float*)st, L0, L1);
  } else {
    launch_translate(c, c->m2l, c->ops.p, c->M0.p, c->M1.p, nullptr, nullptr, c->m2l_stage.p);
    rec(4);
    hipLaunchKernelGGL(k_m2l_reduce<double>, dim3(c->n_boxes), dim3(256), 0, c->stream, c->C, c->m2l_rorder.p, c->m2l_tptr.p,
                       c->m2l_tpairs.p, (const double*)c->m2l_stage.p, L0, L1);
  }
  rec(5);
  // L2L into level l only from parents at levels with M2L pairs (the locals of coarser levels are zero)
  for (int l = std::max(1, c->m2l_min_level + 1); l < pl.n_levels; ++l) {
    launch_translate(c, c->l2l[l], c->ops.p, L0, L1, L0, L1);
  }
  rec(6);
  if (split) HIPCHECK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
  const int part = split ? 2 : 0;
  if (lo) launch_gemv(c, c->near, c->near_mat32.p, c->buf0.p, c->buf1.p, c->y0.p, c->y1.p, part);
  else launch_gemv(c, c->near, c->near.mat.p, c->buf0.p, c->buf1.p, c->y0.p, c->y1.p, part);
  rec(7);
  hipLaunchKernelGGL(k_assemble, dim3(grid1(n)), dim3(256), 0, c->stream, mode, n, c->kind.p, c->point.p, c->alpha.p, phi,
                     c->y0.p, c->y1.p, c->phi_k.p, out);
  rec(8);
  HIPCHECK(hipGetLastError());
}

void drop_graphs(bemfmm_ctx* c) {
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
    evaluate_launch(c, mode, in, out, lo);
    return;
  }
  const int gi = lo ? 1 : 0;
  const size_t bytes = (size_t)c->n_dof * sizeof(double);
  if (!c->gexec[gi]) {
    hipGraph_t g = nullptr;
    HIPCHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    try {
      evaluate_launch(c, 0, c->gin.p, c->gout.p, lo);
    } catch (...) {
      (void)hipStreamEndCapture(c->stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIPCHECK(hipStreamEndCapture(c->stream, &g));
    const hipError_t e = hipGraphInstantiate(&c->gexec[gi], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHECK(e);
  }
  if (in != c->gin.p && in != held)
    HIPCHECK(hipMemcpyAsync(c->gin.p, in, bytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCHECK(hipGraphLaunch(c->gexec[gi], c->stream));
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
  for (int32_t d : c->face_dof)
    if (d < 0 || d >= n) throw std::runtime_error("bemfmm_prepare: face_dof out of range");
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
    DBuf<double> dX;
    DBuf<int32_t> dpt, v0, v1;
    DBuf<uint64_t> k0, k1;
    dX.upload(c->X, s);
    dpt.upload(c->dof_point, s);
    v0.alloc(n);
    v1.alloc(n);
    k0.alloc(n);
    k1.alloc(n);
    hipLaunchKernelGGL(k_morton_keys, dim3((n + 255) / 256), dim3(256), 0, s, n, dpt.p, dX.p, lo[0], hi[0], lo[1], hi[1],
                       lo[2], hi[2], k0.p, v0.p);
    HIPCHECK(hipGetLastError());
    size_t tmp_bytes = 0;
    HIPCHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, k0.p, k1.p, v0.p, v1.p, n, 0, 63, s));
    DBuf<unsigned char> tmp;
    tmp.alloc(tmp_bytes);
    HIPCHECK(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, k0.p, k1.p, v0.p, v1.p, n, 0, 63, s));
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
  for (size_t k = 0; k < c->face_dof.size(); ++k) c->face_dof_i[k] = c->inv[c->face_dof[k]];
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
std::vector<int32_t> update_operator_table(bemfmm_ctx* c, const Plan& pl) {
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
    const auto key = bemfmm::op_key(pl, op);
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
  // upload the new operators in batches through pinned staging, padded to [KP][RP]
  constexpr int kBatch = 32;
  if (!fresh.empty() && c->op_stage_n < (size_t)kBatch * per) {
    if (c->op_stage) (void)hipHostFree(c->op_stage);
    c->op_stage = nullptr;
    HIPCHECK(hipHostMalloc((void**)&c->op_stage, (size_t)kBatch * per * sizeof(double), hipHostMallocDefault));
    c->op_stage_n = (size_t)kBatch * per;
  }
  for (size_t b0 = 0; b0 < fresh.size(); b0 += kBatch) {
    const int nb = (int)std::min<size_t>(kBatch, fresh.size() - b0);
    for (int i = 0; i < nb; ++i) {
      const auto key = bemfmm::op_key(pl, fresh[b0 + i]);
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

// Rotation factors of the plan's M2L offsets in the FP32 table rot32 [slot][RS], indexed by the operator
// table's slots; only slots without factors yet are computed (process-wide cache) and uploaded.
void update_rotation_table(bemfmm_ctx* c, const Plan& pl, const std::vector<int32_t>& slot) {
  hipStream_t s = c->stream;
  const int RS = rot_size(pl.P);  // padded device layout
  if (c->rot_RS != RS) {
    c->rot32.release();
    c->rot_have.clear();
    c->rot_RS = RS;
  }
  if ((int64_t)c->rot_have.size() < c->ops_cap) {
    DBuf<float> t;
    t.alloc((size_t)c->ops_cap * RS);
    if (!c->rot_have.empty())
      HIPCHECK(hipMemcpyAsync(t.p, c->rot32.p, c->rot_have.size() * RS * sizeof(float), hipMemcpyDeviceToDevice, s));
    HIPCHECK(hipStreamSynchronize(s));
    c->rot32 = std::move(t);
    c->rot_have.resize(c->ops_cap, 0);
  }
  std::vector<int> fresh;
  for (int op = 16; op < pl.n_ops; ++op)
    if (!c->rot_have[slot[op]]) {
      c->rot_have[slot[op]] = 1;
      fresh.push_back(op);
    }
  if (fresh.empty()) return;
  {
    std::vector<std::array<int, 4>> keys;
    for (int op : fresh) {
      const auto k = bemfmm::op_key(pl, op);
      keys.push_back({3, k[1], k[2], k[3]});
    }
    bemfmm::warm_operator_cache(pl.P, keys);  // factors of new offsets, computed in parallel once per process
  }
  const size_t cap = c->op_stage_n * 2 / RS;  // rotations per pinned staging batch (the staging holds doubles)
  if (cap < 1) {
    if (c->op_stage) (void)hipHostFree(c->op_stage);
    c->op_stage = nullptr;
    c->op_stage_n = (size_t)RS * 64;
    HIPCHECK(hipHostMalloc((void**)&c->op_stage, c->op_stage_n * sizeof(double), hipHostMallocDefault));
  }
  const size_t batch = c->op_stage_n * 2 / RS;
  float* stg = reinterpret_cast<float*>(c->op_stage);
  for (size_t b0 = 0; b0 < fresh.size(); b0 += batch) {
    const size_t nb = std::min(batch, fresh.size() - b0);
    // consecutive slots are uploaded together: runs of fresh ops with consecutive slots
    for (size_t i = 0; i < nb; ++i) {
      const auto key = bemfmm::op_key(pl, fresh[b0 + i]);
      const auto v = bemfmm::cached_operator(pl.P, 3, key[1], key[2], key[3]);
      rot_pad(pl.P, v->data(), stg + i * RS);
    }
    size_t i = 0;
    while (i < nb) {
      size_t j = i + 1;
      while (j < nb && slot[fresh[b0 + j]] == slot[fresh[b0 + j - 1]] + 1) ++j;
      HIPCHECK(hipMemcpyAsync(c->rot32.p + (size_t)slot[fresh[b0 + i]] * RS, stg + i * RS, (j - i) * RS * sizeof(float),
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
    DTiles alt;
    alt.upload(pl.m2l, c->stream, c->m2l_slot, known);
    c->m2l = std::move(alt);
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
  DTiles alt;
  alt.upload(pl.m2l, s, c->m2l_slot, c->m2l.tp == 16 ? 32 : 16);
  const float t_cur = time_tiles(c->m2l), t_alt = time_tiles(alt);
  if (t_alt < t_cur) c->m2l = std::move(alt);
  c->m2l_tp = c->m2l.tp;
  {
    std::lock_guard<std::mutex> lk(mu);
    chosen.emplace(cls, c->m2l_tp);
  }
  HIPCHECK(hipGetLastError());
}

void do_prepare(bemfmm_ctx* c) {
  if (!c->have_geom) throw std::runtime_error("bemfmm_prepare: geometry not set");
  if (!c->have_bc) throw std::runtime_error("bemfmm_prepare: boundary values not set");
  if (c->opts.precision != BEMFMM_PRECISION_FP64 && c->opts.precision != BEMFMM_PRECISION_FP32_REFINE)
    throw std::runtime_error("bemfmm_prepare: unknown precision mode");
  if (c->opts.m2l_tile_pairs != 0 && c->opts.m2l_tile_pairs != 16 && c->opts.m2l_tile_pairs != 32 &&
      c->opts.m2l_tile_pairs != 64)
    throw std::runtime_error("bemfmm_prepare: m2l_tile_pairs must be 0 (default), 16, 32 or 64");
  if (c->opts.m2l_kernel < 0 || c->opts.m2l_kernel > 2)
    throw std::runtime_error("bemfmm_prepare: m2l_kernel must be 0 (default), 1 (matrix cores) or 2 (rotation)");
  if (c->opts.m2l_kernel == 2 && !rot_compiled(c->opts.p))
    throw std::runtime_error("bemfmm_prepare: the rotation M2L is compiled for p = 8 and 10 only");
  if (c->opts.preconditioner == BEMFMM_PRECOND_SCHWARZ &&
      (c->opts.schwarz_max_core < 1 || c->opts.schwarz_max_block < c->opts.schwarz_max_core ||
       c->opts.schwarz_max_block > kSchwarzMaxBlock || c->opts.schwarz_overlap < 0 || c->opts.schwarz_overlap > 8))
    throw std::runtime_error("bemfmm_prepare: Schwarz block sizes need 1 <= max_core <= max_block <= 128, 0 <= overlap <= 8");
  const double t0 = now();
  bemfmm::Options o;
  o.p = c->opts.p;
  o.theta = c->opts.theta;
  o.leaf_max = c->opts.leaf_max;
  o.max_level = c->opts.max_level;
  o.q = c->opts.q;
  o.materialize_ops = false;  // operators go from the float128 cache straight into the device table
  o.fill_p2m = false;         // the P2M matrices are built on the GPU (k_p2m_build)
  c->prepared = false;
  drop_graphs(c);  // their kernels point into the buffers reallocated below
  // the Krylov basis is sized for n_dof: drop it (re-grown by the next solve)
  c->V.release();
  c->partial.release();
  c->hbuf.release();
  c->v_cap = 0;
  dof_order(c);
  c->plan.reset(new Plan());
  Plan& pl = *c->plan;
  bemfmm::build_plan(pl, o, c->n_nodes, c->X.data(), c->n_faces, c->tri.data(), c->face_dof_i.data(), c->n_dof,
                     c->kind_i.data(), c->point_i.data());
  c->C = pl.C;
  c->n_boxes = pl.n_boxes;
  c->p2m_rows_pad = pl.p2m_rows_pad;
  if (pl.p2m_rows_pad > 4 * 128) throw std::runtime_error("expansion order too large for the P2M kernel");
  if ((pl.C + 63) / 64 > 7) throw std::runtime_error("expansion order too large for the translation kernel");
  if (2 * pl.C > 256) throw std::runtime_error("expansion order too large for the M2L reduction");
  hipStream_t s = c->stream;
  const double t1 = now();
  // ---- upload ----
  c->near.upload(pl.near, s);
  c->near.mat.alloc(pl.near.mat_size);
  c->p2m.upload(pl.p2m, s);
  c->p2m.mat.alloc(pl.p2m.mat_size);
  fill(c, c->p2m.mat.p, pl.p2m.mat_size, 0.);  // rows past C stay zero
  c->RP = (pl.C + 63) / 64 * 64;
  c->KP = (pl.C + 4 * kKG - 1) / (4 * kKG) * (4 * kKG);
  const std::vector<int32_t> slot = update_operator_table(c, pl);
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
    const size_t col = std::max<size_t>(pl.C, (stage_stride<float>(pl.C) + 1) / 2);
    c->m2l_stage.alloc(std::max<size_t>(1, pl.m2l.dst.size() * 2 * col));
  }
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
  DBuf<int32_t> row_block, ncol_dof, face_perm, fv, fdof;
  DBuf<int64_t> range_ptr, ranges, range_prefix;
  DBuf<double> Xs, fX, fcent, fcr, fJ, l2p, d_dun;
  {
    std::vector<int32_t> rb(pl.near.row_out.size());
    for (int b = 0; b < pl.near.n_blocks(); ++b)
      for (int64_t r = pl.near.row_ptr[b]; r < pl.near.row_ptr[b + 1]; ++r) rb[r] = b;
    row_block.upload(rb, s);
  }
  ncol_dof.upload(pl.near_ncol_dof, s);
  c->near_ncd.upload(pl.near_ncol_dof, s);
  face_perm.upload(pl.face_perm, s);
  fv.upload(pl.fv, s);
  fdof.upload(pl.fdof, s);
  range_ptr.upload(pl.near_range_ptr, s);
  ranges.upload(pl.near_ranges, s);
  range_prefix.upload(pl.near_range_prefix, s);
  Xs.upload(pl.Xs, s);
  fX.upload(pl.fX, s);
  fcent.upload(pl.fcent, s);
  fcr.upload(pl.fcr, s);
  fJ.upload(pl.fJ, s);
  l2p.upload(pl.l2p_coef, s);
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
  }
  // ---- P2M matrices on the GPU ----
  if (c->p2m.nblk > 0) {
    if ((int)pl.p2m_face_ptr.size() != c->p2m.nblk + 1 || pl.p2m_face_lc.size() != 3 * pl.p2m_faces.size())
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
    DBuf<int64_t> fptr;
    DBuf<int32_t> pfaces, plc;
    DBuf<double> pcs, dcnm;
    fptr.upload(pl.p2m_face_ptr, s);
    pfaces.upload(pl.p2m_faces, s);
    plc.upload(pl.p2m_face_lc, s);
    pcs.upload(pl.p2m_leaf_cs, s);
    dcnm.upload(cnm, s);
#define BEMFMM_P2MB(PV)                                                                                              \
  hipLaunchKernelGGL(k_p2m_build<PV>, dim3(c->p2m.nblk), dim3(256), 0, s, o.q, pl.p2m_rows_pad, fptr.p, pfaces.p, \
                     plc.p, pcs.p, c->p2m.mat_ptr.p, c->p2m.col_ptr.p, fX.p, fcr.p, fJ.p, d_dun.p, dcnm.p, c->p2m.mat.p)
    switch (pl.P) {
    case 1: BEMFMM_P2MB(1); break;
    case 2: BEMFMM_P2MB(2); break;
    case 3: BEMFMM_P2MB(3); break;
    case 4: BEMFMM_P2MB(4); break;
    case 5: BEMFMM_P2MB(5); break;
    case 6: BEMFMM_P2MB(6); break;
    case 7: BEMFMM_P2MB(7); break;
    case 8: BEMFMM_P2MB(8); break;
    case 9: BEMFMM_P2MB(9); break;
    case 10: BEMFMM_P2MB(10); break;
    default: throw std::runtime_error("expansion order outside 1..10 for the P2M build");
    }
#undef BEMFMM_P2MB
    HIPCHECK(hipGetLastError());
  }
  HIPCHECK(hipStreamSynchronize(s));
  const double t2 = now();
  // ---- near-field quadrature on the GPU ----
  fill(c, c->near.mat.p, pl.near.mat_size, 0.);
  const int64_t nrows = (int64_t)pl.near.row_out.size();
  if (nrows > 0)
    hipLaunchKernelGGL(k_near_quad, dim3((unsigned)nrows), dim3(256), 0, s, pl.C, o.q, pl.near_region, row_block.p,
                       c->near.row_ptr.p, c->near.col_ptr.p, c->near.mat_ptr.p, c->near.row_out.p, c->near.col_idx.p,
                       ncol_dof.p, range_ptr.p, ranges.p, range_prefix.p, face_perm.p, Xs.p, fX.p, fcent.p, fcr.p,
                       fJ.p, fv.p, fdof.p, l2p.p, d_dun.p, c->near.mat.p);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(s));
  const double t3 = now();
  // ---- rigid-mode alpha (updateGeometricProperties) ----
  evaluate(c, 2, nullptr, c->alpha.p);
  tune_m2l_tiles(c, pl);
  const int pc = c->opts.preconditioner;
  c->sw_host = bemfmm::SchwarzBlocks{};
  for (auto* b : {&c->sw_inv}) b->release();
  if (pc == BEMFMM_PRECOND_JACOBI || pc == BEMFMM_PRECOND_SCHWARZ) {
    std::vector<int64_t> node_row(c->n_nodes, -1);
    for (size_t r = 0; r < pl.near.row_out.size(); ++r) node_row[pl.near.row_out[r]] = (int64_t)r;
    for (int i = 0; i < c->n_dof; ++i)
      if (c->kind_i[i] != 2 && node_row[c->point_i[i]] < 0)
        throw std::runtime_error("preconditioner: a DOF's point has no near-field row");
    DBuf<int64_t> d_node_row;
    d_node_row.upload(node_row, s);
    if (pc == BEMFMM_PRECOND_JACOBI) {
      hipLaunchKernelGGL(k_diag_inv, dim3(grid1(c->n_dof)), dim3(256), 0, s, c->n_dof, c->kind.p, c->robin.p, c->point.p,
                         d_node_row.p,
                         row_block.p, c->near.row_ptr.p, c->near.col_ptr.p, c->near.mat_ptr.p, c->near.col_idx.p,
                         ncol_dof.p, c->near.mat.p, c->alpha.p, c->jacobi.p);
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
      bemfmm::build_schwarz_blocks(c->n_dof, c->n_faces, c->face_dof_i.data(), c->point_i.data(), seed,
                                   c->opts.schwarz_max_core, c->opts.schwarz_max_block, c->opts.schwarz_overlap,
                                   c->sw_host);
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
      DBuf<int32_t> failed;
      failed.alloc(nb);
      const size_t lds = (size_t)B.max_block * B.max_block * sizeof(double) + (size_t)B.max_block * 16;
      hipLaunchKernelGGL(k_schwarz_build, dim3(nb), dim3(256), lds, s, c->sw_ptr.p, c->sw_idx.p, c->sw_ncore.p,
                         c->sw_invptr.p, c->kind.p, c->robin.p, c->point.p, d_node_row.p, row_block.p, c->near.row_ptr.p,
                         c->near.col_ptr.p, c->near.mat_ptr.p, c->near.col_idx.p, ncol_dof.p, c->near.mat.p, c->alpha.p,
                         c->sw_inv.p, failed.p);
      HIPCHECK(hipGetLastError());
      std::vector<int32_t> fh(nb);
      HIPCHECK(hipMemcpyAsync(fh.data(), failed.p, nb * sizeof(int32_t), hipMemcpyDeviceToHost, s));
      HIPCHECK(hipStreamSynchronize(s));
      c->sw_failed = 0;
      for (int32_t f : fh) c->sw_failed += f;
    }
  } else if (pc != BEMFMM_PRECOND_NONE) {
    throw std::runtime_error("unknown preconditioner");
  }
  if (c->opts.precision == BEMFMM_PRECISION_FP32_REFINE) {
    // round the resident FP64 matrices once to FP32 (operators were computed in binary128 and rounded to FP64)
    c->near_mat32.alloc(pl.near.mat_size);
    hipLaunchKernelGGL(k_to_f32, dim3(grid1(pl.near.mat_size)), dim3(256), 0, s, (int64_t)pl.near.mat_size, c->near.mat.p,
                       c->near_mat32.p);
    c->p2m_mat32.alloc(c->p2m.mat.n);
    hipLaunchKernelGGL(k_to_f32, dim3(grid1(c->p2m.mat.n)), dim3(256), 0, s, (int64_t)c->p2m.mat.n, c->p2m.mat.p,
                       c->p2m_mat32.p);
    c->m2l_rot = c->opts.m2l_kernel != 1 && rot_compiled(pl.P);
    if (c->m2l_rot) {
      update_rotation_table(c, pl, c->m2l_slot);
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
        hipLaunchKernelGGL(k_to_f32, dim3(grid1(n32)), dim3(256), 0, s, n32, c->ops.p + c->ops32_valid * per,
                           c->ops32.p + c->ops32_valid * per);
        c->ops32_valid = c->ops_used;
      }
    }
    HIPCHECK(hipGetLastError());
  } else {
    c->near_mat32.release();
    c->p2m_mat32.release();
    c->ops32.release();
    c->ops32_valid = 0;
    c->m2l_rot = false;
  }
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
                             c->near_mat32.bytes() + c->p2m_mat32.bytes() + c->ops32.bytes() + c->sw_inv.bytes());
  st.t_tree = pl.t_tree;
  st.t_lists = pl.t_lists;
  st.t_near_struct = pl.t_near_struct;
  st.t_p2m = pl.t_p2m;
  st.t_ops = pl.t_ops;
  st.t_upload = t2 - t1;
  st.t_near_quad = t3 - t2;
  st.t_alpha = t4 - t3;
  st.n_precond_blocks = c->sw_host.n_blocks();
  st.precond_fallback_blocks = c->sw_failed;
  st.precond_avg_block = c->sw_host.n_blocks() ? (double)c->sw_host.idx.size() / c->sw_host.n_blocks() : 0.;
  st.n_ops_uploaded = c->ops_uploaded;
  st.m2l_tile_pairs = c->m2l_tp;
  st.m2l_rotation = c->m2l_rot ? 1 : 0;
  (void)t0;
}

void ensure_basis(bemfmm_ctx* c, int64_t need) {
  if (need <= c->v_cap) return;
  int64_t cap = std::max<int64_t>(need, std::max<int64_t>(64, 2 * c->v_cap));
  DBuf<double> nv;
  nv.alloc((size_t)cap * c->n_dof);
  if (c->v_cap > 0)
    HIPCHECK(hipMemcpyAsync(nv.p, c->V.p, (size_t)c->v_cap * c->n_dof * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  std::swap(c->V.p, nv.p);
  std::swap(c->V.n, nv.n);
  c->v_cap = cap;
  c->partial.alloc((size_t)kDotBlocks * (cap + 1));
  // [h1 (k) | h2 (k) | w.w] of one Arnoldi step (k <= cap), or the k update coefficients
  c->hbuf.alloc(2 * (size_t)cap + 2);
  if (c->hres) (void)hipHostFree(c->hres);
  c->hres = nullptr;
  HIPCHECK(hipHostMalloc((void**)&c->hres, (2 * (size_t)cap + 2) * sizeof(double), hipHostMallocDefault));
}

// hbuf[off, off + k) = V[0..k) . w on the device (no host copy)
void dots_dev(bemfmm_ctx* c, int k, const double* Vp, const double* w, size_t off) {
  const int64_t n = c->n_dof;
  // the reduction buffers are sized by ensure_basis for up to v_cap + 1 products
  if (k < 1 || k > c->v_cap + 1 || !c->partial.p || !c->hbuf.p || (size_t)k * kDotBlocks > c->partial.n ||
      off + (size_t)k > c->hbuf.n)
    throw HipError("internal: dot-product buffers not sized for " + std::to_string(k) + " products", BEMFMM_E_STATE);
  hipLaunchKernelGGL(k_dots, dim3(kDotBlocks, k), dim3(256), 0, c->stream, Vp, n, w, n, c->partial.p);
  hipLaunchKernelGGL(k_reduce_partials, dim3(k), dim3(256), 0, c->stream, c->partial.p, kDotBlocks, k, c->hbuf.p + off);
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
  const int gN = grid1(n);
  switch (c->opts.preconditioner) {
  case BEMFMM_PRECOND_JACOBI:
    if (in != out) hipLaunchKernelGGL(k_mul_to, dim3(gN), dim3(256), 0, c->stream, n, c->jacobi.p, in, out);
    else hipLaunchKernelGGL(k_mul, dim3(gN), dim3(256), 0, c->stream, n, c->jacobi.p, out);
    break;
  case BEMFMM_PRECOND_SCHWARZ:
    if (in == out) throw std::runtime_error("internal: Schwarz apply needs distinct buffers");
    hipLaunchKernelGGL(k_schwarz_apply, dim3(c->sw_host.n_blocks()), dim3(256), 0, c->stream, c->sw_ptr.p, c->sw_idx.p,
                       c->sw_ncore.p, c->sw_invptr.p, c->sw_inv.p, in, out);
    break;
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
    hipLaunchKernelGGL(k_scale_dup, dim3(grid1(n)), dim3(256), 0, c->stream, n, a, w, c->gin.p);
    c->gin_holds = w;
  } else {
    hipLaunchKernelGGL(k_axpby, dim3(grid1(n)), dim3(256), 0, c->stream, n, a, w, 0.0, w);
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
  const int gN = grid1(n);
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
      hipLaunchKernelGGL(k_sub, dim3(gN), dim3(256), 0, s, n, R.rhs, c->tmpv.p, c->V.p);
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
      hipLaunchKernelGGL(k_vsum, dim3(gN), dim3(256), 0, s, c->V.p, n, nb, c->hbuf.p, -1.0, w, n);
      dots_dev(c, nb, c->V.p, w, nb);
      hipLaunchKernelGGL(k_vsum, dim3(gN), dim3(256), 0, s, c->V.p, n, nb, c->hbuf.p + nb, -1.0, w, n);
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
      hipLaunchKernelGGL(k_vsum, dim3(gN), dim3(256), 0, s, c->V.p, n, (int)k, c->hbuf.p, 1.0, R.x, n);
      HIPCHECK(hipStreamSynchronize(s));
    };
    auto true_residual = [&]() {
      Av(R.x, c->tmpv.p);
      hipLaunchKernelGGL(k_sub, dim3(gN), dim3(256), 0, s, n, c->tmpv.p, R.rhs, c->tmpv.p);
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

// FP64 preconditioned residual r = M^-1 (b - A x) into out; returns |r|_2.
double fp64_residual(bemfmm_ctx* c, const double* x, double* out) {
  const int64_t n = c->n_dof;
  const int gN = grid1(n);
  apply_op(c, x, out, false);
  hipLaunchKernelGGL(k_sub, dim3(gN), dim3(256), 0, c->stream, n, c->bvec.p, out, out);
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
  const int gN = grid1(n);
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
      fill(c, c->dvec.p, n, 0.);
      GmresRun in{c->rvec.p, c->dvec.p, std::max(eta * cur_err, 0.5 * tol), max_iter, restart, true, false, true};
      gmres_run(c, in, iters);
      hipLaunchKernelGGL(k_axpby, dim3(gN), dim3(256), 0, s, n, 1.0, c->dvec.p, 1.0, c->xsol.p);
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
  hipLaunchKernelGGL(k_perm_gather, dim3((n + 255) / 256), dim3(256), 0, c->stream, n, c->d_perm.p, x, c->xin.p);
  evaluate(c, 0, c->xin.p, c->yout.p, lo);
  hipLaunchKernelGGL(k_perm_scatter, dim3((n + 255) / 256), dim3(256), 0, c->stream, n, c->d_perm.p, c->yout.p, y);
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
  static char name[256] = {0};
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
  o->use_graphs = 1;
  o->m2l_kernel = 0;
  o->overlap_near = 1;
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
    HIPCHECK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    return BEMFMM_OK;
  });
  if (rc != BEMFMM_OK) {
    std::fprintf(stderr, "bemfmm_ctx_create: %s\n", c->err.c_str());
    delete c;
    return nullptr;
  }
  return c;
}

void bemfmm_ctx_destroy(bemfmm_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->stream2) (void)hipStreamSynchronize(c->stream2);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto e : {c->ev_fork, c->ev_join})
    if (e) (void)hipEventDestroy(e);
  drop_graphs(c);
  if (c->op_stage) (void)hipHostFree(c->op_stage);
  if (c->hres) (void)hipHostFree(c->hres);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
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
    c->tri.assign(tri, tri + 3 * (size_t)n_faces);
    c->face_dof.assign(face_dof, face_dof + 3 * (size_t)n_faces);
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
