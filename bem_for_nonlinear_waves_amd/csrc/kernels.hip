// MI355X (gfx950) device kernels of the FMM-accelerated GMRES BVP engine and their launchers
// (kernels.hpp). The host side (bemfmm.cpp) calls only the launchers.
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
// Setup: the near-field quadrature (GW5x5 singular/near, Dunavant far; GMRES_FMM.hpp:1824-1891), the P2M
// matrices, the Morton keys and the Schwarz block inverses run here too.
#include <hip/hip_runtime.h>

#include <cstring>  // rocPRIM's texture iterator uses memset without including it
#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <utility>

#include "kernels.hpp"
#include "nearquad.hpp"

namespace {

using namespace bemfmm::dev;

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
// inputs are staged through LDS in chunks; each wavefront streams RPW rows at once with
// 16-byte vector loads (double2 / float4: rows are padded to a multiple of 4 columns and start
// 16-byte aligned) and issues all 2*RPW*2 loads of two column steps before using them (rows
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
// RPW: rows per wave and pass (each staged x value is read from LDS once per RPW rows); chosen per problem class
// by the autotuner of the prepare (2, 4 or 8)
template <typename T, int RPW>
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
    for (int rg = wave * RPW; rg < nrow; rg += 4 * RPW) {
      const T* Ar[RPW];
#pragma unroll
      for (int k = 0; k < RPW; ++k) Ar[k] = M + (int64_t)min(rg + k, nrow - 1) * 2 * ncol + cs;
      double s0[RPW], s1[RPW];
#pragma unroll
      for (int k = 0; k < RPW; ++k) s0[k] = s1[k] = 0.;
      int c = lane * W;
      for (; c + STEP < cn; c += 2 * STEP) {
        V va[RPW][2], vb[RPW][2];
#pragma unroll
        for (int k = 0; k < RPW; ++k) {
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
        for (int k = 0; k < RPW; ++k)
#pragma unroll
          for (int u = 0; u < W; ++u) {
            s0[k] += GemvVec<T>::get(va[k][0], u) * xa[0][u] + GemvVec<T>::get(va[k][1], u) * xa[1][u];
            s1[k] += GemvVec<T>::get(vb[k][0], u) * xb[0][u] + GemvVec<T>::get(vb[k][1], u) * xb[1][u];
          }
      }
      if (c < cn) {
        V va[RPW], vb[RPW];
#pragma unroll
        for (int k = 0; k < RPW; ++k) {
          va[k] = ld_nt(reinterpret_cast<const V*>(Ar[k] + c));
          vb[k] = ld_nt(reinterpret_cast<const V*>(Ar[k] + ncol + c));
        }
#pragma unroll
        for (int k = 0; k < RPW; ++k)
#pragma unroll
          for (int u = 0; u < W; ++u) {
            s0[k] += GemvVec<T>::get(va[k], u) * xs0[c + u];
            s1[k] += GemvVec<T>::get(vb[k], u) * xs1[c + u];
          }
      }
#pragma unroll
      for (int k = 0; k < RPW; ++k) {
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

// The FP32 operator's near field folded for A x (mode 0, the only one the FP32 operator runs): in A x a DOF
// column meets either phi_n = x (Dirichlet: weight G) or phi = x (Neumann and CORNER: -Gn; ROBIN: phi = x and
// phi_n = kappa x, kappa G - Gn), so one weight per entry carries the row's y0 - y1. Block b of the FP64 near field
// (rows x [G | Gn], columns [DOF (ncd) | L2P (C) | padding]) becomes rows x U, columns [DOF (ncd) | L2P on the
// single-layer locals (Cl) | -L2P on the double-layer locals (Cl) | padding to 4]; the column indices address
// [x (n) | L0 (nbC = boxes x Cl) | L1 (nbC)]. One workgroup per block.
// Cl <= C: the locals' coefficients per box, and the L2P columns kept per layer (the inner order's (p_in+1)^2 when
// the FP32 operator runs its expansions at a lower order; else C).
__global__ __launch_bounds__(256) void k_fold_near(int C, int Cl, int n, int64_t nbC, const int64_t* __restrict__ row_ptr,
                                                   const int64_t* __restrict__ col_ptr, const int64_t* __restrict__ mat_ptr,
                                                   const int32_t* __restrict__ ncol_dof,
                                                   const int32_t* __restrict__ col_idx, const double* __restrict__ mat,
                                                   const int32_t* __restrict__ kind, const double* __restrict__ kappa,
                                                   const int64_t* __restrict__ ucol_ptr,
                                                   const int64_t* __restrict__ umat_ptr, int32_t* __restrict__ ucol_idx,
                                                   float* __restrict__ umat) {
  const int b = blockIdx.x;
  const int nrow = (int)(row_ptr[b + 1] - row_ptr[b]);
  const int ncol = (int)(col_ptr[b + 1] - col_ptr[b]);
  const int ncu = (int)(ucol_ptr[b + 1] - ucol_ptr[b]);
  const int ncd = ncol_dof[b];
  const int32_t* ci = col_idx + col_ptr[b];
  int32_t* cu = ucol_idx + ucol_ptr[b];
  // L2P columns: local j of the target box, at stride Cl in [L0 | L1] (the FP64 field's index is n + box C + j)
  const int32_t l0 = n + (ci[ncd] - n) / C * Cl;
  for (int c = threadIdx.x; c < ncu; c += blockDim.x)
    cu[c] = c < ncd ? ci[c] : (c < ncd + Cl ? l0 + (c - ncd) : (c < ncd + 2 * Cl ? l0 + (int32_t)nbC + (c - ncd - Cl) : l0));
  const double* A = mat + mat_ptr[b];
  float* U = umat + umat_ptr[b];
  for (int64_t e = threadIdx.x; e < (int64_t)nrow * ncu; e += blockDim.x) {
    const int r = (int)(e / ncu), c = (int)(e - (int64_t)r * ncu);
    const double* G = A + (int64_t)r * 2 * ncol;
    const double* Gn = G + ncol;
    double u = 0.;
    if (c < ncd) {
      const int d = ci[c];
      const int k = kind[d];
      u = k == 0 ? G[c] : (k == 3 ? kappa[d] * G[c] - Gn[c] : -Gn[c]);
    } else if (c < ncd + Cl) {
      u = G[c];
    } else if (c < ncd + 2 * Cl) {
      u = -Gn[c - Cl];
    }
    U[(int64_t)r * ncu + c] = (float)u;
  }
}

// y = U x over the folded FP32 near field (k_fold_near): the block GEMV of k_block_gemv with one weight per entry
// and one sum per row; the staged inputs are gathered from x (column < n) or the locals L0 / L1. part as in
// k_block_gemv (1: the DOF columns, written; 2: the rest, added).
template <int RPW>
__global__ __launch_bounds__(256) void k_block_gemv_u(const int64_t* __restrict__ row_ptr, const int64_t* __restrict__ col_ptr,
                                                      const int64_t* __restrict__ mat_ptr, const int32_t* __restrict__ row_out,
                                                      const int32_t* __restrict__ col_idx, const float* __restrict__ mat,
                                                      int n, int64_t nbC, const double* __restrict__ x,
                                                      const double* __restrict__ L0, const double* __restrict__ L1,
                                                      double* __restrict__ y, int part, const int32_t* __restrict__ ncol_dof) {
  typedef float4 V;
  constexpr int W = 4;
  constexpr int STEP = 64 * W;
  __shared__ double xs[kGemvChunk];
  __shared__ double acc[kGemvMaxRows];
  const int b = blockIdx.x;
  const int64_t r0 = row_ptr[b];
  const int nrow = (int)(row_ptr[b + 1] - r0);
  const int64_t c0 = col_ptr[b];
  const int ncol = (int)(col_ptr[b + 1] - c0);  // multiple of 4
  const float* M = mat + mat_ptr[b];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < nrow; i += blockDim.x) acc[i] = 0.;
  const int split = part == 0 ? 0 : (ncol_dof[b] & ~3);
  const int cb = part == 2 ? split : 0, ce = part == 1 ? split : ncol;
  auto gather = [&](int g) -> double { return g < n ? x[g] : (g < n + nbC ? L0[g - n] : L1[g - n - nbC]); };
  for (int cs = cb; cs < ce; cs += kGemvChunk) {
    const int cn = min(kGemvChunk, ce - cs);  // multiple of 4
    __syncthreads();
    const int32_t* ci = col_idx + c0 + cs;
    int i = threadIdx.x;
    for (; i + 3 * 256 < cn; i += 4 * 256) {
      const int g0 = ci[i], g1 = ci[i + 256], g2 = ci[i + 512], g3 = ci[i + 768];
      const double a0 = gather(g0), a1 = gather(g1), a2 = gather(g2), a3 = gather(g3);
      xs[i] = a0;
      xs[i + 256] = a1;
      xs[i + 512] = a2;
      xs[i + 768] = a3;
    }
    for (; i < cn; i += 256) xs[i] = gather(ci[i]);
    __syncthreads();
    for (int rg = wave * RPW; rg < nrow; rg += 4 * RPW) {
      const float* Ar[RPW];
#pragma unroll
      for (int k = 0; k < RPW; ++k) Ar[k] = M + (int64_t)min(rg + k, nrow - 1) * ncol + cs;
      double s0[RPW];
#pragma unroll
      for (int k = 0; k < RPW; ++k) s0[k] = 0.;
      int c = lane * W;
      for (; c + STEP < cn; c += 2 * STEP) {
        V va[RPW][2];
#pragma unroll
        for (int k = 0; k < RPW; ++k) {
          va[k][0] = ld_nt(reinterpret_cast<const V*>(Ar[k] + c));
          va[k][1] = ld_nt(reinterpret_cast<const V*>(Ar[k] + c + STEP));
        }
        double xa[2][W];
#pragma unroll
        for (int u = 0; u < W; ++u) {
          xa[0][u] = xs[c + u];
          xa[1][u] = xs[c + STEP + u];
        }
#pragma unroll
        for (int k = 0; k < RPW; ++k)
#pragma unroll
          for (int u = 0; u < W; ++u)
            s0[k] += GemvVec<float>::get(va[k][0], u) * xa[0][u] + GemvVec<float>::get(va[k][1], u) * xa[1][u];
      }
      if (c < cn) {
        V va[RPW];
#pragma unroll
        for (int k = 0; k < RPW; ++k) va[k] = ld_nt(reinterpret_cast<const V*>(Ar[k] + c));
#pragma unroll
        for (int k = 0; k < RPW; ++k)
#pragma unroll
          for (int u = 0; u < W; ++u) s0[k] += GemvVec<float>::get(va[k], u) * xs[c + u];
      }
#pragma unroll
      for (int k = 0; k < RPW; ++k) {
        double u = s0[k];
        for (int off = 32; off > 0; off >>= 1) u += __shfl_down(u, off);
        if (lane == 0 && rg + k < nrow) acc[rg + k] += u;
      }
    }
  }
  __syncthreads();
  for (int r = threadIdx.x; r < nrow; r += blockDim.x) {
    const int o = row_out[r0 + r];
    if (part == 2) y[o] += acc[r];  // each row belongs to one block: a plain read-modify-write
    else y[o] = acc[r];
  }
}

// P2M: one workgroup per leaf; the leaf's matrix is column-major per layer (plan.hpp), so thread
// (layer = tid / 128, row = tid % 128 + 128 k) streams a column's rows with its neighbours: every load is
// coalesced and every lane has a row, however few DOF columns the leaf has. The gathered phi_n / phi
// values are staged in LDS in chunks. M[leaf][row] = sum_col A[layer][col][row] x_layer[col] (FP64 sums).
// U: column loads in flight per thread (4, 8 or 16), chosen per problem class by the autotuner of the prepare
constexpr int kP2MChunk = 1024;
template <typename T, int U>
__global__ __launch_bounds__(256) void k_p2m(int C, int CP, int Crows, int Cout, const int64_t* __restrict__ row_ptr,
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
      if (r >= Crows) break;
      const T* a = Al + (int64_t)cs * CP + r;
      double s0 = 0., s1 = 0.;
      int c = 0;
      for (; c + U - 1 < cn; c += U) {  // U loads in flight (streamed: non-temporal)
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = (double)__builtin_nontemporal_load(&a[(int64_t)(c + u) * CP]);
#pragma unroll
        for (int u = 0; u < U; u += 2) {
          s0 += v[u] * xl[c + u];
          s1 += v[u + 1] * xl[c + u + 1];
        }
      }
      for (; c < cn; ++c) s0 += (double)__builtin_nontemporal_load(&a[(int64_t)c * CP]) * xl[c];
      acc[k] += s0 + s1;
    }
  }
  const int64_t base = (int64_t)(row_out[row_ptr[b]] / C) * Cout;  // leaf box * Cout (row_out holds leaf * C)
  double* out = layer ? M1 : M0;
#pragma unroll
  for (int k = 0; k < kRows; ++k) {
    const int r = r0 + 128 * k;
    if (r < Crows) out[base + r] = acc[k];
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
// NPF = 3: linear triangles (fX 9 values per face, one normal); NPF = 6: true quadratic faces (fX 18 values, the
// point and normal from the 6-node geometry at each Dunavant point, densities weighted by the redistributed
// TriShape<6>; a face-local column of -1 marks an edge midpoint without a DOF).
template <int P, int NPF>
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
  __shared__ double pw[kP2MPts][NPF], pn[kP2MPts][3];
  __shared__ int plc[kP2MPts][NPF];
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
      double x, y, z;
      if constexpr (NPF == 3) {
        const double J = fJ[f];
        x = (b0 * fX[9 * f] + b1 * fX[9 * f + 3] + b2 * fX[9 * f + 6] - cx) / s;
        y = (b0 * fX[9 * f + 1] + b1 * fX[9 * f + 4] + b2 * fX[9 * f + 7] - cy) / s;
        z = (b0 * fX[9 * f + 2] + b1 * fX[9 * f + 5] + b2 * fX[9 * f + 8] - cz) / s;
        const double W = J * w;
        pw[tid][0] = W * b0;
        pw[tid][1] = W * b1;
        pw[tid][2] = W * b2;
        for (int a = 0; a < 3; ++a) {
          plc[tid][a] = face_lc[3 * (f0 + fi) + a];
          pn[tid][a] = fcr[3 * f + a] / J;
        }
      } else {
        double xq[3], cr[3], N[6];
        bemfmm::tri6_geom(fX + 18 * (int64_t)f, b0, b1, xq, cr);
        unsigned active = 0;
        for (int e = 0; e < 3; ++e) active |= (face_lc[6 * (f0 + fi) + 3 + e] >= 0 ? 1u : 0u) << e;
        bemfmm::tri6_density(b0, b1, active, N);
        const double Jq = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        x = (xq[0] - cx) / s;
        y = (xq[1] - cy) / s;
        z = (xq[2] - cz) / s;
        const double W = Jq * w;
        for (int k = 0; k < 6; ++k) {
          pw[tid][k] = W * N[k];
          plc[tid][k] = face_lc[6 * (f0 + fi) + k];
        }
        for (int a = 0; a < 3; ++a) pn[tid][a] = Jq > 1e-30 ? cr[a] / Jq : fcr[3 * f + a] / fJ[f];
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
        for (int k = 0; k < NPF; ++k)
          if (NPF == 3 || plc[pt][k] >= 0) M[(int64_t)(layer * ncol + plc[pt][k]) * CP + row] += pw[pt][k] * v;
      }
    }
    __syncthreads();
  }
}

// FP32 multipoles for the rotation M2L: M32[(box*2 + layer)*C + r] = M_layer[box*Cs + r], r < C: the first C
// coefficients (degrees <= P of a P' >= P expansion stored with stride Cs: the packing is degree-major)
__global__ void k_pack_m32(int64_t n, int C, int Cs, const double* __restrict__ M0, const double* __restrict__ M1,
                           float* __restrict__ M32) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t col = i / C, r = i - col * C;
    M32[i] = (float)((col & 1) ? M1 : M0)[(col >> 1) * Cs + r];
  }
}

template <int P>
__global__ __launch_bounds__(256) void k_m2l_rot(int ntiles, const float* __restrict__ rot,
                                                 const int32_t* __restrict__ tile_op, const int64_t* __restrict__ tile_p0,
                                                 const int64_t* __restrict__ tile_p1, const int32_t* __restrict__ src,
                                                 const float* __restrict__ M32, float* __restrict__ stage) {
  constexpr int C = (P + 1) * (P + 1);
  constexpr int RS = rot_size(P);
  constexpr int NCH = stage_stride<float>(C) / 16;  // 16-row chunks of the staged column (rows >= C are zero)
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

// FP32 staging: the same reduction with 16-byte loads. A pair's two staged columns (CS = stage_stride<float>(C)
// floats each) are LP = CS / 2 float4s; a wave reads 64 / LP pairs at once (1 at C = 121, 2 at the inner orders
// 5 and 6, whose columns are 64 floats), lane (sub, t) float4 t of its pair. The workgroup's G = 4 x 64 / LP
// lane groups take the pairs a + g, a + g + G, ... (kReduceLoads in flight per lane, summed in pair order) and
// the groups' sums are combined in group order (deterministic). Boxes run in descending pair count (order),
// as in k_m2l_reduce.
constexpr int kReduceLoads = 8;
// Cout >= C: the locals are stored with stride Cout; rows [C, Cout) (degrees above the M2L's order) are zeroed.
__global__ __launch_bounds__(256) void k_m2l_reduce_f32(int C, int Cout, const int32_t* __restrict__ order,
                                                        const int64_t* __restrict__ tptr,
                                                        const int32_t* __restrict__ tpairs,
                                                        const float* __restrict__ stage, double* __restrict__ L0,
                                                        double* __restrict__ L1) {
  __shared__ double part[4][256];  // [component][group g * LP + t]
  const int CS = stage_stride<float>(C);  // <= 128, a multiple of 32
  const int LP = CS / 2, ppw = 64 / LP, G = 4 * ppw;
  const int tb = order[blockIdx.x];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int sub = lane / LP, t = lane - sub * LP;
  const bool act = sub < ppw;  // lanes past ppw * LP idle (LP = 48 at C = 81)
  const int g = w * ppw + sub;
  const int64_t a = tptr[tb], b = tptr[tb + 1];
  double s0 = 0., s1 = 0., s2 = 0., s3 = 0.;
  if (act) {
    int64_t q = a + g;
    for (; q + (int64_t)G * (kReduceLoads - 1) < b; q += (int64_t)G * kReduceLoads) {
      float4 v[kReduceLoads];
#pragma unroll
      for (int u = 0; u < kReduceLoads; ++u)
        v[u] = ld_nt(reinterpret_cast<const float4*>(stage + (int64_t)tpairs[q + (int64_t)G * u] * 2 * CS) + t);
#pragma unroll
      for (int u = 0; u < kReduceLoads; ++u) {
        s0 += v[u].x;
        s1 += v[u].y;
        s2 += v[u].z;
        s3 += v[u].w;
      }
    }
    for (; q < b; q += G) {
      const float4 v = ld_nt(reinterpret_cast<const float4*>(stage + (int64_t)tpairs[q] * 2 * CS) + t);
      s0 += v.x;
      s1 += v.y;
      s2 += v.z;
      s3 += v.w;
    }
    part[0][g * LP + t] = s0;
    part[1][g * LP + t] = s1;
    part[2][g * LP + t] = s2;
    part[3][g * LP + t] = s3;
  }
  for (int r = C + threadIdx.x; r < Cout; r += blockDim.x) {
    L0[(int64_t)tb * Cout + r] = 0.;
    L1[(int64_t)tb * Cout + r] = 0.;
  }
  __syncthreads();
  if (threadIdx.x < LP) {
    const int tt = threadIdx.x;
    double S0 = 0., S1 = 0., S2 = 0., S3 = 0.;
    for (int gg = 0; gg < G; ++gg) {
      S0 += part[0][gg * LP + tt];
      S1 += part[1][gg * LP + tt];
      S2 += part[2][gg * LP + tt];
      S3 += part[3][gg * LP + tt];
    }
    const int f = 4 * tt;  // a float4 never straddles the two layers (CS is a multiple of 4)
    const int layer = f >= CS ? 1 : 0, r = f - layer * CS;
    double* L = (layer ? L1 : L0) + (int64_t)tb * Cout;
    if (r < C) L[r] = S0;
    if (r + 1 < C) L[r + 1] = S1;
    if (r + 2 < C) L[r + 2] = S2;
    if (r + 3 < C) L[r + 3] = S3;
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
  // y1 == nullptr: y0 already holds y0 - y1 (the folded FP32 near field of A x)
  double v = (kind[i] == 2) ? phi[i] : y0[p] - ((y1 ? y1[p] : 0.) + alpha[i] * phi[i]);
  if (mode == 1) v = (kind[i] == 2) ? phi_k[i] : -v;
  out[i] = v;
}

// Near-field quadrature: one workgroup per near row (target node); threads stride over the
// faces of the target box's near ranges and scatter (WG, WGn) of the three corners into the row.
// fill_direct_linear / fill_nonadj_linear (BEM_solveBVP_GMRES_FMM.hpp:1824-1891).
template <int NPF>
__global__ __launch_bounds__(256) void k_near_quad(
    int C, int q, double near_region, const int32_t* __restrict__ row_block, const int64_t* __restrict__ row_ptr,
    const int64_t* __restrict__ col_ptr, const int64_t* __restrict__ mat_ptr, const int32_t* __restrict__ row_out,
    const int32_t* __restrict__ col_idx, const int32_t* __restrict__ ncol_dof, const int64_t* __restrict__ range_ptr,
    const int64_t* __restrict__ ranges, const int64_t* __restrict__ range_prefix, const int32_t* __restrict__ face_perm,
    const double* __restrict__ Xs, const double* __restrict__ fX, const double* __restrict__ fcent,
    const double* __restrict__ fcr, const double* __restrict__ fJ, const int32_t* __restrict__ fv,
    const int32_t* __restrict__ fdof, const double* __restrict__ l2p, const double* __restrict__ dun_rule,
    const double* __restrict__ dun6_rule, const int32_t* __restrict__ face_cols, const int64_t* __restrict__ block_fstart, int lds_cols,
    double* __restrict__ mat) {
  // a row of at most lds_cols DOF columns accumulates in LDS (ds_add_f64) and is written once; a longer one
  // adds into the zero-filled matrix with global atomics
  extern __shared__ __align__(16) double acc[];  // 2 * lds_cols
  // Dunavant rule of the non-adjacent, non-near faces ({b0,b1,b2,w} x q), per context
  __shared__ double dun[7][4], dun6[6][4];
  for (int i = threadIdx.x; i < 28; i += blockDim.x) dun[i / 4][i % 4] = dun_rule[i];
  if (NPF == 6)
    for (int i = threadIdx.x; i < 24; i += blockDim.x) dun6[i / 4][i % 4] = dun6_rule[i];
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
  const bool in_lds = ncd <= lds_cols;
  double* S0 = in_lds ? acc : A0;
  double* S1 = in_lds ? acc + ncd : A1;
  if (in_lds) {
    for (int c = threadIdx.x; c < 2 * ncd; c += blockDim.x) acc[c] = 0.;
    __syncthreads();
  }
  const double x = Xs[3 * t], y = Xs[3 * t + 1], z = Xs[3 * t + 2];
  const int64_t g0 = range_ptr[b], g1 = range_ptr[b + 1];
  const int64_t nf = range_prefix[g1 - 1] + (ranges[2 * (g1 - 1) + 1] - ranges[2 * (g1 - 1)]);
  const int32_t* cols = col_idx + c0;
  const int32_t* fcols = face_cols ? face_cols + NPF * block_fstart[b] : nullptr;
  for (int64_t i = threadIdx.x; i < nf; i += blockDim.x) {
    // locate the range holding local face i
    int64_t lo = g0, hi = g1 - 1;
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (range_prefix[mid] <= i) lo = mid;
      else hi = mid - 1;
    }
    const int64_t f = ranges[2 * lo] + (i - range_prefix[lo]);  // face arrays are in face_perm order
    int slot = -1;
    for (int k = 0; k < NPF; ++k)
      if (fv[NPF * f + k] == t) slot = k;
    double WG[NPF], WGn[NPF];
    if constexpr (NPF == 3) {
      bemfmm::face_weights(fX + 9 * f, fJ[f], fcr + 3 * f, fcent + 3 * f, slot, x, y, z, near_region, c_gw5, dun, q, WG,
                           WGn);
    } else {
      unsigned active = 0;
      for (int e = 0; e < 3; ++e) active |= (fdof[6 * f + 3 + e] >= 0 ? 1u : 0u) << e;
      bemfmm::face_weights_quad(fX + 18 * f, active, fcent + 3 * f, slot, x, y, z, near_region, c_gw5, dun, q, dun6,
                                WG, WGn);
    }
#pragma unroll
    for (int k = 0; k < NPF; ++k) {
      const int d = fdof[NPF * f + k];
      if (NPF == 6 && d < 0) continue;  // edge midpoint without a DOF: its weights were redistributed
      int lo2 = 0;
      if (fcols) {
        lo2 = fcols[NPF * i + k];
      } else {
        int hi2 = ncd - 1;
        while (lo2 < hi2) {
          const int mid = (lo2 + hi2) >> 1;
          if (cols[mid] < d) lo2 = mid + 1;
          else hi2 = mid;
        }
      }
      atomicAdd(&S0[lo2], WG[k]);
      atomicAdd(&S1[lo2], WGn[k]);
    }
  }
  if (in_lds) {
    __syncthreads();
    for (int c = threadIdx.x; c < ncd; c += blockDim.x) {
      A0[c] = acc[c];
      A1[c] = acc[ncd + c];
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

// entry sources of the block build: the resident near field, or the rows of the reference's sparse matrix.
// row(i) resolves once what every entry of row i needs (its sorted column list and its values), col(r, k) is
// the k-th column of that list and val(r, k, i, j) the operator entry when column j sits at position k; the
// block build searches the columns of several rows at once, so that their loads overlap.
struct SrcRow {
  int64_t base = 0;  // start of the sorted column list
  int64_t aux = 0;   // start of the row's values
  int n = 0;         // columns in the list; -1: identity row (CORNER)
  int ncol = 0;      // near field: stride between the G and Gn rows
};
struct NearSrc {
  NearRows a;
  __device__ SrcRow row(int i) const {
    SrcRow r;
    if (a.kind[i] == 2) {
      r.n = -1;
      return r;
    }
    const int64_t q = a.node_row[a.point[i]];
    const int b = a.row_block[q];
    r.base = a.col_ptr[b];
    r.ncol = (int)(a.col_ptr[b + 1] - r.base);
    r.n = a.ncol_dof[b];
    r.aux = a.mat_ptr[b] + (q - a.row_ptr[b]) * 2 * r.ncol;
    return r;
  }
  __device__ int col(const SrcRow& r, int k) const { return a.col_idx[r.base + k]; }
  __device__ double val(const SrcRow& r, int k, int i, int j) const {
    const double g = a.mat[r.aux + k];
    double v = (a.kind[j] == 0) ? g : -a.mat[r.aux + r.ncol + k];
    if (a.kind[j] == 3) v += a.kappa[j] * g;
    (void)i;
    return v;
  }
  // Neumann-type diagonals carry -alpha (also where the column search found nothing)
  __device__ double diag_shift(int i) const {
    const int ki = a.kind[i];
    return (ki == 1 || ki == 3) ? -a.alpha[i] : 0.0;
  }
  __device__ double operator()(int i, int j) const {
    return near_entry(i, j, a.kind, a.kappa, a.point, a.node_row, a.row_block, a.row_ptr, a.col_ptr, a.mat_ptr,
                      a.col_idx, a.ncol_dof, a.mat, a.alpha);
  }
};
struct RowsSrc {
  IluRows A;
  __device__ SrcRow row(int i) const {
    SrcRow r;
    r.base = A.ptr[i];
    r.aux = r.base;
    r.n = (int)(A.ptr[i + 1] - r.base);
    return r;
  }
  __device__ int col(const SrcRow& r, int k) const { return A.col[r.base + k]; }
  __device__ double val(const SrcRow& r, int k, int, int) const { return A.val[r.aux + k]; }
  __device__ double diag_shift(int) const { return 0.0; }
  __device__ double operator()(int i, int j) const {
    int64_t lo = A.ptr[i], hi = A.ptr[i + 1];  // columns ascending: first column >= j
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (A.col[mid] < j) lo = mid + 1;
      else hi = mid;
    }
    return (lo < A.ptr[i + 1] && A.col[lo] == j) ? A.val[lo] : 0.0;
  }
};

// One workgroup per block (m <= 128 DOFs): extract A_b into LDS, invert it in place by Gauss-Jordan
// elimination with partial pivoting, and store the core rows of A_b^-1 (ncore x m, row-major).
// A block with a vanishing or non-finite pivot falls back to the inverse diagonal on its core
// (as SchwarzAdditive_CRS does when its block LU fails) and is flagged in failed[b].
// LDS: m*m doubles (matrix) + m doubles (pivot column) + m ints (DOF ids) + m ints (pivot rows) + m row
// references. The gather takes one column per thread and searches it in kGatherRows rows at once (their
// dependent loads overlap instead of following each other).
// rows_out: the stored rows per block (the core for the restricted combination, all m for the averaged one).
// 1024 threads (16 waves) per workgroup: the block's LDS allows one workgroup per CU, and the gather's loads and
// the elimination's LDS reads need that many waves in flight (256 threads: 140 ms for the 28,632 blocks of the
// 209k tank, latency-bound).
constexpr int kSchwarzThreads = 1024;
static_assert(kSchwarzMaxBlock <= 128, "k_schwarz_build updates one column per thread, 8 rows per pass");
template <class Src>
__global__ __launch_bounds__(kSchwarzThreads) void k_schwarz_build(
    const int64_t* __restrict__ bptr, const int32_t* __restrict__ bidx, const int32_t* __restrict__ rows_out,
    const int64_t* __restrict__ inv_ptr, Src entry, double* __restrict__ inv, int32_t* __restrict__ failed) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int b = blockIdx.x;
  const int64_t p0 = bptr[b];
  const int m = (int)(bptr[b + 1] - p0);
  const int nc = rows_out[b];
  double* A = reinterpret_cast<double*>(smem);   // m*m
  double* colk = A + m * m;                      // m
  int* ids = reinterpret_cast<int*>(colk + m);   // m
  int* piv = ids + m;                            // m
  SrcRow* rows = reinterpret_cast<SrcRow*>(piv + m + (m & 1));  // m, 8-byte aligned
  __shared__ int bad;
  const int tid = threadIdx.x;
  if (tid == 0) bad = 0;
  for (int t = tid; t < m; t += blockDim.x) ids[t] = bidx[p0 + t];
  __syncthreads();
  for (int t = tid; t < m; t += blockDim.x) rows[t] = entry.row(ids[t]);
  __syncthreads();
  {
    constexpr int kGatherRows = 8;
    const int j = tid & 127;
    if (j < m) {
      const int dj = ids[j];
      constexpr int kRowGroups = kSchwarzThreads / 128;
      for (int i0 = tid >> 7; i0 < m; i0 += kRowGroups * kGatherRows) {
        int lo[kGatherRows], hi[kGatherRows];
#pragma unroll
        for (int u = 0; u < kGatherRows; ++u) {
          const int i = i0 + kRowGroups * u;
          lo[u] = 0;
          hi[u] = (i < m && rows[i].n > 0) ? rows[i].n - 1 : 0;
        }
        bool active = true;
        while (active) {
          active = false;
#pragma unroll
          for (int u = 0; u < kGatherRows; ++u)
            if (lo[u] < hi[u]) {
              const int mid = (lo[u] + hi[u]) >> 1;
              if (entry.col(rows[i0 + kRowGroups * u], mid) < dj) lo[u] = mid + 1;
              else hi[u] = mid;
              active = true;
            }
        }
#pragma unroll
        for (int u = 0; u < kGatherRows; ++u) {
          const int i = i0 + kRowGroups * u;
          if (i >= m) continue;
          const SrcRow& r = rows[i];
          double v;
          if (r.n < 0) {
            v = (i == j) ? 1.0 : 0.0;
          } else {
            v = (r.n > 0 && entry.col(r, lo[u]) == dj) ? entry.val(r, lo[u], ids[i], dj) : 0.0;
            if (i == j) v += entry.diag_shift(ids[i]);
          }
          A[i * m + j] = v;
        }
      }
    }
  }
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
    {  // one column per thread, 8 rows per pass (m <= 128 = blockDim.x / 8): consecutive threads take
       // consecutive columns of a row (no bank conflicts) and no thread divides by m
      const int j = tid & 127;
      if (j < m) {
        const double akj = A[k * m + j];  // the scaled pivot row (row k is not updated here)
        for (int i = tid >> 7; i < m; i += kSchwarzThreads / 128) {
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
        const double d = entry(ids[i], ids[i]);
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

// z[core DOFs of b] = (A_b^-1 r|_b)[core]: one wavefront per block (4 blocks per workgroup), the block's
// gathered residual (m <= kSchwarzMaxBlock = 128 values) held two per lane, the core rows taken four at a time
// so that their loads are in flight together; each DOF is written by exactly the one block whose core holds it
// (restricted additive Schwarz), so z needs no clearing or atomics. (A workgroup per block left most of its
// waves idle on the 8-row cores: 84 us per apply at 209k nodes.)
__global__ __launch_bounds__(256) void k_schwarz_apply(int nblk, const int64_t* __restrict__ bptr,
                                                       const int32_t* __restrict__ bidx,
                                                       const int32_t* __restrict__ bncore, const int64_t* __restrict__ inv_ptr,
                                                       const double* __restrict__ inv, const double* __restrict__ r,
                                                       double* __restrict__ z) {
  static_assert(kSchwarzMaxBlock <= 128, "two residual values per lane");
  const int lane = threadIdx.x & 63;
  const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (b >= nblk) return;  // no workgroup barrier below
  const int64_t p0 = bptr[b];
  const int m = (int)(bptr[b + 1] - p0);
  const int nc = bncore[b];
  const bool h0 = lane < m, h1 = lane + 64 < m;
  const double r0 = h0 ? r[bidx[p0 + lane]] : 0.;
  const double r1 = h1 ? r[bidx[p0 + lane + 64]] : 0.;
  const double* Ib = inv + inv_ptr[b];
  for (int i0 = 0; i0 < nc; i0 += 4) {
    double s[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double* row = Ib + (int64_t)min(i0 + u, nc - 1) * m;
      s[u] = h0 ? row[lane] * r0 : 0.;
      if (h1) s[u] += row[lane + 64] * r1;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
      for (int u = 0; u < 4; ++u) s[u] += __shfl_down(s[u], off);
    if (lane == 0)
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (i0 + u < nc) z[bidx[p0 + i0 + u]] = s[u];
  }
}

// Averaged additive Schwarz (SchwarzAdditive_CRS::solve, basic_linear_systems.hpp:2960-3003), two passes
// with no atomics: every block's local solution A_b^-1 r|_b (all m rows) into y at the block's positions, then
// each DOF sums the entries of the blocks that hold it (in block order) and divides by their number.
__global__ __launch_bounds__(256) void k_schwarz_solve_all(const int64_t* __restrict__ bptr,
                                                           const int32_t* __restrict__ bidx,
                                                           const int64_t* __restrict__ inv_ptr,
                                                           const double* __restrict__ inv,
                                                           const double* __restrict__ r, double* __restrict__ y) {
  __shared__ double rl[kSchwarzMaxBlock];
  const int b = blockIdx.x;
  const int64_t p0 = bptr[b];
  const int m = (int)(bptr[b + 1] - p0);
  for (int t = threadIdx.x; t < m; t += blockDim.x) rl[t] = r[bidx[p0 + t]];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const double* Ib = inv + inv_ptr[b];
  for (int i = wave; i < m; i += 4) {
    const double* row = Ib + (int64_t)i * m;
    double s = 0.;
    for (int j = lane; j < m; j += 64) s += row[j] * rl[j];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
    if (lane == 0) y[p0 + i] = s;
  }
}
__global__ void k_schwarz_average(int n, const int64_t* __restrict__ occ_ptr, const int64_t* __restrict__ occ_pos,
                                  const double* __restrict__ y, double* __restrict__ z) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const int64_t a = occ_ptr[g], e = occ_ptr[g + 1];
  double s = 0.;
  for (int64_t q = a; q < e; ++q) s += y[occ_pos[q]];
  z[g] = e > a ? s / (double)(e - a) : 0.;
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
// ILU(0) sweeps: one row per thread, forward z_i -= sum_{q < diag} L_iq z_col, backward
// z_i = (z_i - sum_{q > diag} U_iq z_col) / U_ii. The columns a row reads belong to earlier levels.
__device__ __forceinline__ void ilu_row(const IluRows& A, int32_t i, bool backward, double* z) {
  const int64_t d = A.diag[i];
  double s = z[i];
  if (!backward) {
    for (int64_t q = A.ptr[i]; q < d; ++q) s -= A.val[q] * z[A.col[q]];
    z[i] = s;
  } else {
    for (int64_t q = d + 1; q < A.ptr[i + 1]; ++q) s -= A.val[q] * z[A.col[q]];
    z[i] = s / A.val[d];
  }
}
__global__ __launch_bounds__(256) void k_ilu_level(IluRows A, const int32_t* __restrict__ rows, int64_t nrows,
                                                   bool backward, double* z) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t < nrows) ilu_row(A, rows[t], backward, z);
}
// a run of small levels in one workgroup: the barrier (with its workgroup-scope fence) makes each level's
// stores visible to the next level's loads
__global__ __launch_bounds__(kIluRunThreads) void k_ilu_run(IluRows A, const int32_t* __restrict__ rows,
                                                            const int32_t* __restrict__ lptr, int l0, int l1,
                                                            bool backward, double* z) {
  for (int l = l0; l < l1; ++l) {
    const int a = lptr[l], b = lptr[l + 1];
    for (int r = a + (int)threadIdx.x; r < b; r += blockDim.x) ilu_row(A, rows[r], backward, z);
    __threadfence_block();
    __syncthreads();
  }
}

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
// out[i] = sum_j P[i * nk + j] in[j] for i < nk, out[i] = in[i] for nk <= i < nb (one workgroup; each sum in
// column order): the Gram-Schmidt coefficients over kept basis vectors projected onto a kept subspace
__global__ void k_proj_coef(int nk, int nb, const double* __restrict__ P, const double* __restrict__ in,
                            double* __restrict__ out) {
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    double v = in[i];
    if (i < nk) {
      v = 0.;
      for (int j = 0; j < nk; ++j) v += P[(int64_t)i * nk + j] * in[j];
    }
    out[i] = v;
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
// Work memory of the allocation-free entry points (bemfmm_matvec_device_noalloc): its address is host-side
// only a stand-in, so host code uses the device address that scratch_region looks up.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

int grid1(int64_t n, int bs = 256) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + bs - 1) / bs, 65536)); }

template <typename T>
void launch_translate(hipStream_t s, int C, int RP, int KP, const Tiles& t, const T* ops, const double* S0,
                      const double* S1, double* D0, double* D1, T* stage) {
  if (t.ntiles == 0) return;
  const size_t lds = translate_lds(t.tp, KP, RP) * sizeof(T);
  const int grid = ((t.ntiles + kXCD - 1) / kXCD) * kXCD;
  const int RTW = RP / 64;
#define BEMFMM_TR_TP(N, TPV)                                                                                         \
  if (stage)                                                                                                         \
    hipLaunchKernelGGL((k_translate_mfma<T, N, true, TPV>), dim3(grid), dim3(256), lds, s, C, RP, KP, t.ntiles, ops, \
                       t.op, t.p0, t.p1, t.dst, t.src, S0, S1, D0, D1, stage);                                       \
  else                                                                                                               \
    hipLaunchKernelGGL((k_translate_mfma<T, N, false, TPV>), dim3(grid), dim3(256), lds, s, C, RP, KP, t.ntiles,     \
                       ops, t.op, t.p0, t.p1, t.dst, t.src, S0, S1, D0, D1, stage)
#define BEMFMM_TR(N)                                   \
  do {                                                 \
    if (!stage && t.tp == 8) BEMFMM_TR_TP(N, 8);       \
    else if (t.tp == 16) BEMFMM_TR_TP(N, 16);          \
    else if (stage && t.tp == 64) BEMFMM_TR_TP(N, 64); \
    else BEMFMM_TR_TP(N, 32);                          \
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

}  // namespace

namespace bemfmm {
namespace dev {

void fill(hipStream_t s, double* p, int64_t n, double v) {
  hipLaunchKernelGGL(k_fill, dim3(grid1(n)), dim3(256), 0, s, p, n, v);
}

void scatter(hipStream_t s, int mode, int n, const int32_t* kind, const double* kappa, const double* x,
             const double* phi_k, const double* phin_k, double* phin, double* phi) {
  hipLaunchKernelGGL(k_scatter, dim3(grid1(n)), dim3(256), 0, s, mode, n, kind, kappa, x, phi_k, phin_k, phin, phi);
}

template <typename T>
void block_gemv_t(hipStream_t s, const Gemv& g, const T* mat, const double* x0, const double* x1, double* y0,
                  double* y1, int part) {
  if (g.nblk == 0) return;
  auto k = g.variant == 1 ? k_block_gemv<T, 2> : g.variant == 2 ? k_block_gemv<T, 8> : k_block_gemv<T, 4>;
  hipLaunchKernelGGL(k, dim3(g.nblk), dim3(256), 0, s, g.row_ptr, g.col_ptr, g.mat_ptr, g.row_out, g.col_idx, mat, x0,
                     x1, y0, y1, part, g.ncol_dof);
}
void block_gemv(hipStream_t s, const Gemv& g, const double* mat, const double* x0, const double* x1, double* y0,
                double* y1, int part) {
  block_gemv_t(s, g, mat, x0, x1, y0, y1, part);
}
void block_gemv(hipStream_t s, const Gemv& g, const float* mat, const double* x0, const double* x1, double* y0,
                double* y1, int part) {
  block_gemv_t(s, g, mat, x0, x1, y0, y1, part);
}

template <typename T>
void p2m_t(hipStream_t s, int C, int CP, int Crows, const Gemv& g, const T* mat, const double* x0, const double* x1,
           double* M0, double* M1) {
  if (g.nblk == 0) return;
  if (Crows < 1 || Crows > C || C > CP) throw std::runtime_error("p2m: rows out of range");
  auto k = g.variant == 1 ? k_p2m<T, 4> : g.variant == 2 ? k_p2m<T, 16> : k_p2m<T, 8>;
  hipLaunchKernelGGL(k, dim3(g.nblk), dim3(256), 0, s, C, CP, Crows, Crows == C ? C : Crows, g.row_ptr, g.col_ptr,
                     g.mat_ptr, g.row_out, g.col_idx, mat, x0, x1, M0, M1);
}
void p2m(hipStream_t s, int C, int CP, const Gemv& g, const double* mat, const double* x0, const double* x1,
         double* M0, double* M1, int Crows) {
  p2m_t(s, C, CP, Crows > 0 ? Crows : C, g, mat, x0, x1, M0, M1);
}
void p2m(hipStream_t s, int C, int CP, const Gemv& g, const float* mat, const double* x0, const double* x1,
         double* M0, double* M1, int Crows) {
  p2m_t(s, C, CP, Crows > 0 ? Crows : C, g, mat, x0, x1, M0, M1);
}

void fold_near(hipStream_t s, const Gemv& g, int C, int Cl, int n, int64_t nbC, const double* mat, const int32_t* kind,
               const double* kappa, const int64_t* ucol_ptr, const int64_t* umat_ptr, int32_t* ucol_idx, float* umat) {
  if (g.nblk == 0) return;
  if (Cl < 1 || Cl > C) throw std::runtime_error("fold_near: L2P columns out of range");
  hipLaunchKernelGGL(k_fold_near, dim3(g.nblk), dim3(256), 0, s, C, Cl, n, nbC, g.row_ptr, g.col_ptr, g.mat_ptr,
                     g.ncol_dof, g.col_idx, mat, kind, kappa, ucol_ptr, umat_ptr, ucol_idx, umat);
}
void block_gemv_u(hipStream_t s, const Gemv& g, const float* mat, int n, int64_t nbC, const double* x,
                  const double* L0, const double* L1, double* y, int part) {
  if (g.nblk == 0) return;
  auto k = g.variant == 1 ? k_block_gemv_u<2> : g.variant == 2 ? k_block_gemv_u<8> : k_block_gemv_u<4>;
  hipLaunchKernelGGL(k, dim3(g.nblk), dim3(256), 0, s, g.row_ptr, g.col_ptr, g.mat_ptr, g.row_out, g.col_idx, mat, n,
                     nbC, x, L0, L1, y, part, g.ncol_dof);
}

void translate(hipStream_t s, int C, int RP, int KP, const Tiles& t, const double* ops, const double* S0,
               const double* S1, double* D0, double* D1, double* stage) {
  launch_translate<double>(s, C, RP, KP, t, ops, S0, S1, D0, D1, stage);
}
void translate(hipStream_t s, int C, int RP, int KP, const Tiles& t, const float* ops, const double* S0,
               const double* S1, double* D0, double* D1, float* stage) {
  launch_translate<float>(s, C, RP, KP, t, ops, S0, S1, D0, D1, stage);
}

void pack_m32(hipStream_t s, int64_t n, int C, int Cs, const double* M0, const double* M1, float* M32) {
  hipLaunchKernelGGL(k_pack_m32, dim3(grid1(n)), dim3(256), 0, s, n, C, Cs, M0, M1, M32);
}

bool m2l_rot_compiled(int P) { return P == 8 || P == 10; }
bool m2l_rot_inner_compiled(int P) { return P == 5 || P == 6 || P == 8 || P == 10; }

void m2l_rot(hipStream_t s, int P, const Tiles& t, const float* rot, const float* M32, float* stage) {
  if (t.ntiles == 0) return;
  const int grid = ((t.ntiles + kXCD - 1) / kXCD) * kXCD;
  switch (P) {
  case 5:
    hipLaunchKernelGGL(k_m2l_rot<5>, dim3(grid), dim3(256), 0, s, t.ntiles, rot, t.op, t.p0, t.p1, t.src, M32, stage);
    break;
  case 6:
    hipLaunchKernelGGL(k_m2l_rot<6>, dim3(grid), dim3(256), 0, s, t.ntiles, rot, t.op, t.p0, t.p1, t.src, M32, stage);
    break;
  case 8:
    hipLaunchKernelGGL(k_m2l_rot<8>, dim3(grid), dim3(256), 0, s, t.ntiles, rot, t.op, t.p0, t.p1, t.src, M32, stage);
    break;
  case 10:
    hipLaunchKernelGGL(k_m2l_rot<10>, dim3(grid), dim3(256), 0, s, t.ntiles, rot, t.op, t.p0, t.p1, t.src, M32, stage);
    break;
  default: throw std::runtime_error("rotation M2L not compiled for this expansion order");
  }
}

void m2l_reduce(hipStream_t s, int C, int nboxes, const int32_t* order, const int64_t* tptr, const int32_t* tpairs,
                const double* stage, double* L0, double* L1) {
  if (nboxes == 0) return;
  hipLaunchKernelGGL(k_m2l_reduce<double>, dim3(nboxes), dim3(256), 0, s, C, order, tptr, tpairs, stage, L0, L1);
}
void m2l_reduce(hipStream_t s, int C, int Cout, int nboxes, const int32_t* order, const int64_t* tptr,
                const int32_t* tpairs, const float* stage, double* L0, double* L1) {
  if (nboxes == 0) return;
  if (Cout < C || Cout > 256) throw std::runtime_error("m2l_reduce: local stride out of range");
  hipLaunchKernelGGL(k_m2l_reduce_f32, dim3(nboxes), dim3(256), 0, s, C, Cout, order, tptr, tpairs, stage, L0, L1);
}

void assemble(hipStream_t s, int mode, int n, const int32_t* kind, const int32_t* point, const double* alpha,
              const double* phi, const double* y0, const double* y1, const double* phi_k, double* out) {
  hipLaunchKernelGGL(k_assemble, dim3(grid1(n)), dim3(256), 0, s, mode, n, kind, point, alpha, phi, y0, y1, phi_k, out);
}

void morton_keys(hipStream_t s, int n, const int32_t* point, const double* X, const double b[6], uint64_t* keys,
                 int32_t* vals) {
  hipLaunchKernelGGL(k_morton_keys, dim3((n + 255) / 256), dim3(256), 0, s, n, point, X, b[0], b[1], b[2], b[3], b[4],
                     b[5], keys, vals);
}

hipError_t radix_sort_pairs(hipStream_t s, void* tmp, size_t& tmp_bytes, const uint64_t* keys_in, uint64_t* keys_out,
                            const int32_t* vals_in, int32_t* vals_out, int n, int bits) {
  return rocprim::radix_sort_pairs(tmp, tmp_bytes, keys_in, keys_out, vals_in, vals_out, (size_t)n, 0u, (unsigned)bits,
                                   s);
}

void p2m_build(hipStream_t s, int P, const P2MBuild& a, double* mat) {
  if (a.nleaves == 0) return;
#define BEMFMM_P2MB(PV)                                                                                              \
  if (a.npf == 6)                                                                                                    \
    hipLaunchKernelGGL((k_p2m_build<PV, 6>), dim3(a.nleaves), dim3(256), 0, s, a.q, a.CP, a.face_ptr, a.faces,      \
                       a.face_lc, a.leaf_cs, a.mat_ptr, a.col_ptr, a.fX, a.fcr, a.fJ, a.rule, a.cnm, mat);           \
  else                                                                                                               \
    hipLaunchKernelGGL((k_p2m_build<PV, 3>), dim3(a.nleaves), dim3(256), 0, s, a.q, a.CP, a.face_ptr, a.faces,      \
                       a.face_lc, a.leaf_cs, a.mat_ptr, a.col_ptr, a.fX, a.fcr, a.fJ, a.rule, a.cnm, mat)
  switch (P) {
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
}

void near_quad(hipStream_t s, int64_t nrows, const NearQuad& a, const Gemv& g, double* mat) {
  if (nrows == 0) return;
  // rows of up to kNearQuadLdsCols DOF columns (most of them) accumulate in 64 KB of LDS: two workgroups per CU
  const int lds_cols = std::min(a.max_ncd, kNearQuadLdsCols);
  auto kern = a.npf == 6 ? k_near_quad<6> : k_near_quad<3>;
  hipLaunchKernelGGL(kern, dim3((unsigned)nrows), dim3(256), (size_t)2 * lds_cols * sizeof(double), s, a.C, a.q,
                     a.near_region, a.row_block, g.row_ptr, g.col_ptr, g.mat_ptr, g.row_out, g.col_idx, a.ncol_dof,
                     a.range_ptr, a.ranges, a.range_prefix, a.face_perm, a.Xs, a.fX, a.fcent, a.fcr, a.fJ, a.fv, a.fdof,
                     a.l2p, a.dun_rule, a.dun6_rule, a.face_cols, a.block_fstart, lds_cols, mat);
}

void to_f32(hipStream_t s, int64_t n, const double* a, float* b) {
  hipLaunchKernelGGL(k_to_f32, dim3(grid1(n)), dim3(256), 0, s, n, a, b);
}

void diag_inv(hipStream_t s, int n, const NearRows& a, double* dinv) {
  hipLaunchKernelGGL(k_diag_inv, dim3(grid1(n)), dim3(256), 0, s, n, a.kind, a.kappa, a.point, a.node_row, a.row_block,
                     a.row_ptr, a.col_ptr, a.mat_ptr, a.col_idx, a.ncol_dof, a.mat, a.alpha, dinv);
}

static size_t schwarz_build_lds(int m) {
  return (size_t)m * m * sizeof(double) + (size_t)m * sizeof(double) + (size_t)(2 * m + (m & 1)) * sizeof(int) +
         (size_t)m * sizeof(SrcRow);
}

void schwarz_build(hipStream_t s, const Schwarz& b, int max_block, const NearRows& a, double* inv, int32_t* failed) {
  if (b.nblk == 0) return;
  const size_t lds = schwarz_build_lds(max_block);
  hipLaunchKernelGGL(k_schwarz_build<NearSrc>, dim3(b.nblk), dim3(kSchwarzThreads), lds, s, b.bptr, b.bidx, b.bncore, b.inv_ptr,
                     NearSrc{a}, inv, failed);
}
void schwarz_build_rows(hipStream_t s, const Schwarz& b, int max_block, const IluRows& A, const int32_t* rows_out,
                        double* inv, int32_t* failed) {
  if (b.nblk == 0) return;
  const size_t lds = schwarz_build_lds(max_block);
  hipLaunchKernelGGL(k_schwarz_build<RowsSrc>, dim3(b.nblk), dim3(kSchwarzThreads), lds, s, b.bptr, b.bidx, rows_out, b.inv_ptr,
                     RowsSrc{A}, inv, failed);
}
void schwarz_apply_averaged(hipStream_t s, const Schwarz& b, int n, const int64_t* occ_ptr, const int64_t* occ_pos,
                            const double* inv, const double* r, double* y, double* z) {
  if (b.nblk == 0) return;
  hipLaunchKernelGGL(k_schwarz_solve_all, dim3(b.nblk), dim3(256), 0, s, b.bptr, b.bidx, b.inv_ptr, inv, r, y);
  hipLaunchKernelGGL(k_schwarz_average, dim3(grid1(n)), dim3(256), 0, s, n, occ_ptr, occ_pos, y, z);
}

void dots(hipStream_t s, int k, const double* V, int64_t ldv, const double* w, int64_t n, double* partial) {
  hipLaunchKernelGGL(k_dots, dim3(kDotBlocks, k), dim3(256), 0, s, V, ldv, w, n, partial);
}
void reduce_partials(hipStream_t s, const double* partial, int k, double* out) {
  hipLaunchKernelGGL(k_reduce_partials, dim3(k), dim3(256), 0, s, partial, kDotBlocks, k, out);
}
void vsum(hipStream_t s, const double* V, int64_t ldv, int k, const double* h, double sign, double* w, int64_t n) {
  hipLaunchKernelGGL(k_vsum, dim3(grid1(n)), dim3(256), 0, s, V, ldv, k, h, sign, w, n);
}
void proj_coef(hipStream_t s, int nk, int nb, const double* P, const double* in, double* out) {
  if (nb <= 0) return;
  if (nk < 0 || nk > nb || in == out) throw std::runtime_error("proj_coef: invalid sizes or aliased buffers");
  hipLaunchKernelGGL(k_proj_coef, dim3(1), dim3(256), 0, s, nk, nb, P, in, out);
}
void axpby(hipStream_t s, int64_t n, double a, const double* x, double b, double* y) {
  hipLaunchKernelGGL(k_axpby, dim3(grid1(n)), dim3(256), 0, s, n, a, x, b, y);
}
void sub(hipStream_t s, int64_t n, const double* a, const double* b, double* out) {
  hipLaunchKernelGGL(k_sub, dim3(grid1(n)), dim3(256), 0, s, n, a, b, out);
}
void mul(hipStream_t s, int64_t n, const double* d, double* y) {
  hipLaunchKernelGGL(k_mul, dim3(grid1(n)), dim3(256), 0, s, n, d, y);
}
void mul_to(hipStream_t s, int64_t n, const double* d, const double* x, double* y) {
  hipLaunchKernelGGL(k_mul_to, dim3(grid1(n)), dim3(256), 0, s, n, d, x, y);
}
void scale_dup(hipStream_t s, int64_t n, double a, double* w, double* dup) {
  hipLaunchKernelGGL(k_scale_dup, dim3(grid1(n)), dim3(256), 0, s, n, a, w, dup);
}
void schwarz_apply(hipStream_t s, const Schwarz& b, const double* inv, const double* r, double* z) {
  if (b.nblk == 0) return;
  hipLaunchKernelGGL(k_schwarz_apply, dim3((b.nblk + 3) / 4), dim3(256), 0, s, b.nblk, b.bptr, b.bidx, b.bncore,
                     b.inv_ptr, inv, r, z);
}
void ilu_level(hipStream_t s, const IluRows& A, const int32_t* level_rows, int64_t nrows, bool backward, double* z) {
  if (nrows > 0) k_ilu_level<<<(unsigned)((nrows + 255) / 256), 256, 0, s>>>(A, level_rows, nrows, backward, z);
}
void ilu_run(hipStream_t s, const IluRows& A, const int32_t* rows, const int32_t* lptr, int l0, int l1, bool backward,
             double* z) {
  if (l1 > l0) k_ilu_run<<<1, kIluRunThreads, 0, s>>>(A, rows, lptr, l0, l1, backward, z);
}

void perm_gather(hipStream_t s, int n, const int32_t* perm, const double* in, double* out) {
  hipLaunchKernelGGL(k_perm_gather, dim3((n + 255) / 256), dim3(256), 0, s, n, perm, in, out);
}
void perm_scatter(hipStream_t s, int n, const int32_t* perm, const double* in, double* out) {
  hipLaunchKernelGGL(k_perm_scatter, dim3((n + 255) / 256), dim3(256), 0, s, n, perm, in, out);
}

namespace {
std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch` per device (nullptr: not looked up yet)
}  // namespace

hipError_t scratch_region(void** ptr, size_t* bytes) {
  int d = 0;
  hipError_t e = hipGetDevice(&d);
  if (e != hipSuccess) return e;
  if (d < 0 || d >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[d]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[d] = p;
  }
  *ptr = g_scratch_ptr[d];
  *bytes = kScratchBytes;
  return hipSuccess;
}

}  // namespace dev
}  // namespace bemfmm
