// Translation operators of the FMM in real-field packed form, evaluated in
// __float128 and rounded once to double.
//
// Conventions (reference lib/include/lib_multipole_expansion.hpp:578-614, 996-1083;
// lib/include/moments.hpp:165-436, 493-598, 862-884):
//   Y_n^m(theta,phi) = (-1)^|m| sqrt((n-|m|)!/(n+|m|)!) P_n^|m|(cos theta) e^{i m phi}
//   R_n^m(x) = rho^n Y_n^m, I_n^m(x) = Y_n^m / rho^(n+1)
//   multipole  M_n^m = sum q R_n^{-m}(x_s - c);   local field = Re sum L_n^m R_n^m(x - c)
// The reference's translation tables (AAA_{M2M,M2L,L2L}_FMM.hpp) are not in the
// source tree; the coefficients below are Greengard & Rokhlin (1997), Theorems
// 3.3 (M2M), 3.5 (M2L) and 3.6 (L2L), with A_n^m = (-1)^n/sqrt((n-m)!(n+m)!).
//
// Real-field packing: because the boundary data are real, C_n^{-m} = conj(C_n^m)
// (moments.hpp:63-90, 287-298); a coefficient set is stored as (P+1)^2 reals per
// layer: Re C_n^0, then (Re C_n^m, Im C_n^m) for m = 1..n.  A complex operator
// T[(j,k),(n,m)] becomes a real (P+1)^2 x (P+1)^2 matrix acting on that packing.
//
// Level scaling: with box side s, multipoles are stored as M~_n = M_n / s^n and
// locals as L~_j = L_j s^(j+1); every operator is then independent of the level
// (boxes of a level are a regular lattice of pitch s, children have side s/2).
#include <quadmath.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <vector>
#include <stdexcept>

#include "plan.hpp"

namespace bemfmm {
namespace {

typedef __float128 q;

struct qc {
  q re, im;
};
inline qc operator+(qc a, qc b) { return {a.re + b.re, a.im + b.im}; }
inline qc operator-(qc a, qc b) { return {a.re - b.re, a.im - b.im}; }
inline qc operator*(qc a, qc b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline qc operator*(q s, qc a) { return {s * a.re, s * a.im}; }
inline qc conj(qc a) { return {a.re, -a.im}; }

constexpr int kMaxDeg = 48;
struct QTables {
  q fact[2 * kMaxDeg + 2];
  q sqfnm[kMaxDeg + 1][kMaxDeg + 1];  // sqrt((n-m)!(n+m)!), m <= n
  QTables() {
    fact[0] = 1;
    for (int i = 1; i < 2 * kMaxDeg + 2; ++i) fact[i] = fact[i - 1] * i;
    for (int n = 0; n <= kMaxDeg; ++n)
      for (int m = 0; m <= n; ++m) sqfnm[n][m] = sqrtq(fact[n - m] * fact[n + m]);
  }
};
const QTables& qt() {
  static const QTables t;
  return t;
}
q fact(int k) { return qt().fact[k]; }
q Anm(int n, int m) {
  const int a = m < 0 ? -m : m;
  if (a > n) return 0;
  const q v = 1 / qt().sqfnm[n][a];
  return (n & 1) ? -v : v;
}
qc ipow(int e) {
  e = ((e % 4) + 4) % 4;
  switch (e) {
  case 0: return {1, 0};
  case 1: return {0, 1};
  case 2: return {-1, 0};
  default: return {0, -1};
  }
}
inline int iabs(int a) { return a < 0 ? -a : a; }
inline int idx(int n, int m) { return n * (n + 1) + m; }

// Regular solid harmonics R_n^m(x) for all |m| <= n <= L, from the unnormalised
// recurrences NR_m^m = (x+iy)/(2m) NR_{m-1}^{m-1},
// NR_n^m = ((2n-1) z NR_{n-1}^m - r^2 NR_{n-2}^m) / ((n+m)(n-m)), R_n^m = (-1)^m sqrt((n-m)!(n+m)!) NR_n^m.
std::vector<qc> regular(q x, q y, q z, int L) {
  std::vector<qc> R((L + 1) * (L + 1));
  std::vector<qc> NR((L + 1) * (L + 1), qc{0, 0});
  const q r2 = x * x + y * y + z * z;
  NR[idx(0, 0)] = {1, 0};
  for (int m = 1; m <= L; ++m) NR[idx(m, m)] = (1 / (q)(2 * m)) * (qc{x, y} * NR[idx(m - 1, m - 1)]);
  for (int m = 0; m <= L; ++m)
    for (int n = m + 1; n <= L; ++n) {
      qc a = (q)(2 * n - 1) * z * NR[idx(n - 1, m)];
      if (n - 2 >= m) a = a - r2 * NR[idx(n - 2, m)];
      NR[idx(n, m)] = (1 / (q)((n + m) * (n - m))) * a;
    }
  for (int n = 0; n <= L; ++n)
    for (int m = 0; m <= n; ++m) {
      q c = qt().sqfnm[n][m];
      if (m & 1) c = -c;
      R[idx(n, m)] = c * NR[idx(n, m)];
      R[idx(n, -m)] = conj(R[idx(n, m)]);
    }
  return R;
}

// Irregular solid harmonics I_n^m(x) = Y_n^m/r^(n+1) for all |m| <= n <= L, from
// NS_0^0 = 1/r, NS_m^m = (2m-1)(x+iy)/r^2 NS_{m-1}^{m-1},
// NS_n^m = ((2n-1) z NS_{n-1}^m - ((n-1)^2-m^2) NS_{n-2}^m)/r^2, I_n^m = (-1)^m NS_n^m/sqrt((n-m)!(n+m)!).
std::vector<qc> irregular(q x, q y, q z, int L) {
  std::vector<qc> I((L + 1) * (L + 1));
  std::vector<qc> NS((L + 1) * (L + 1), qc{0, 0});
  const q r2 = x * x + y * y + z * z;
  NS[idx(0, 0)] = {1 / sqrtq(r2), 0};
  for (int m = 1; m <= L; ++m) NS[idx(m, m)] = ((q)(2 * m - 1) / r2) * (qc{x, y} * NS[idx(m - 1, m - 1)]);
  for (int m = 0; m <= L; ++m)
    for (int n = m + 1; n <= L; ++n) {
      qc a = (q)(2 * n - 1) * z * NS[idx(n - 1, m)];
      if (n - 2 >= m) a = a - (q)((n - 1) * (n - 1) - m * m) * NS[idx(n - 2, m)];
      NS[idx(n, m)] = (1 / r2) * a;
    }
  for (int n = 0; n <= L; ++n)
    for (int m = 0; m <= n; ++m) {
      q c = 1 / qt().sqfnm[n][m];
      if (m & 1) c = -c;
      I[idx(n, m)] = c * NS[idx(n, m)];
      I[idx(n, -m)] = conj(I[idx(n, m)]);
    }
  return I;
}

// Accumulate the real-field packed form of a complex operator entry
// out_(j,k) += T * in_(n,m) (and the term of in_(n,-m) = conj(in_(n,m))), k >= 0.
struct Packer {
  int P, C;
  std::vector<q> A;  // [k_in][r_out]
  explicit Packer(int P_) : P(P_), C((P_ + 1) * (P_ + 1)), A((size_t)C * C, 0) {}
  // T_pos = T[(j,k),(n,m)], T_neg = T[(j,k),(n,-m)] (m >= 0; for m == 0 pass T_neg = 0)
  void add(int j, int k, int n, int m, qc Tp, qc Tn) {
    const int rr = pk_re(j, k);
    const int ri = k > 0 ? pk_im(j, k) : -1;
    if (m == 0) {
      const int cr = pk_re(n, 0);
      A[(size_t)cr * C + rr] += Tp.re;
      if (ri >= 0) A[(size_t)cr * C + ri] += Tp.im;
      return;
    }
    const qc Sp = Tp + Tn, Sm = Tp - Tn;
    const int cr = pk_re(n, m), ci = pk_im(n, m);
    A[(size_t)cr * C + rr] += Sp.re;
    A[(size_t)ci * C + rr] += -Sm.im;
    if (ri >= 0) {
      A[(size_t)cr * C + ri] += Sp.im;
      A[(size_t)ci * C + ri] += Sm.re;
    }
  }
  void store(double* out) const {
    for (size_t i = 0; i < A.size(); ++i) out[i] = (double)A[i];
  }
};

}  // namespace

void m2l_operator(int P, int dx, int dy, int dz, double* out) {
  // L~_j^k += sum_{n,m} i^{|k-m|-|k|-|m|} A_n^m A_j^k / ((-1)^n A_{j+n}^{m-k}) I_{j+n}^{m-k}(o) M~_n^m,
  // o = (source centre - target centre)/s.
  const std::vector<qc> I = irregular((q)dx, (q)dy, (q)dz, 2 * P);
  Packer pk(P);
  auto T = [&](int j, int k, int n, int m) -> qc {
    const q a = Anm(n, m) * Anm(j, k) / (((n & 1) ? (q)-1 : (q)1) * Anm(j + n, m - k));
    return a * (ipow(iabs(k - m) - iabs(k) - iabs(m)) * I[idx(j + n, m - k)]);
  };
  for (int j = 0; j <= P; ++j)
    for (int k = 0; k <= j; ++k)
      for (int n = 0; n <= P; ++n)
        for (int m = 0; m <= n; ++m) pk.add(j, k, n, m, T(j, k, n, m), m > 0 ? T(j, k, n, -m) : qc{0, 0});
  pk.store(out);
}

void m2m_operator(int P, int octant, double* out) {
  // M~_j^k(parent) += sum_{n,m} i^{|k|-|m|-|k-m|} A_n^m A_{j-n}^{k-m}/A_j^k R_n^{-m}(d) 2^{-j} M~_{j-n}^{k-m}(child),
  // d = (child centre - parent centre)/s_child = (+-1/2, +-1/2, +-1/2).
  const q h = (q)0.5;
  const std::vector<qc> R = regular((octant & 1) ? h : -h, (octant & 2) ? h : -h, (octant & 4) ? h : -h, P);
  Packer pk(P);
  // coefficient of input (n', m') = (j-n, k-m) for output (j,k)
  auto T = [&](int j, int k, int np, int mp) -> qc {
    const int n = j - np, m = k - mp;
    if (n < 0 || iabs(m) > n || iabs(mp) > np) return {0, 0};
    const q a = Anm(n, m) * Anm(np, mp) / Anm(j, k) * powq(2, -j);
    return a * (ipow(iabs(k) - iabs(m) - iabs(k - m)) * R[idx(n, -m)]);
  };
  for (int j = 0; j <= P; ++j)
    for (int k = 0; k <= j; ++k)
      for (int np = 0; np <= j; ++np)
        for (int mp = 0; mp <= np; ++mp) pk.add(j, k, np, mp, T(j, k, np, mp), mp > 0 ? T(j, k, np, -mp) : qc{0, 0});
  pk.store(out);
}

void l2l_operator(int P, int octant, double* out) {
  // L~_j^k(child) += sum_{n>=j,m} i^{|m|-|m-k|-|k|} A_{n-j}^{m-k} A_j^k/((-1)^{n+j} A_n^m) R_{n-j}^{m-k}(d) 2^{-(n+1)} L~_n^m(parent),
  // d = (parent centre - child centre)/s_child.
  const q h = (q)0.5;
  const std::vector<qc> R = regular((octant & 1) ? -h : h, (octant & 2) ? -h : h, (octant & 4) ? -h : h, P);
  Packer pk(P);
  auto T = [&](int j, int k, int n, int m) -> qc {
    if (n < j || iabs(m - k) > n - j || iabs(m) > n) return {0, 0};
    const q a = Anm(n - j, m - k) * Anm(j, k) / ((((n + j) & 1) ? (q)-1 : (q)1) * Anm(n, m)) * powq(2, -(n + 1));
    return a * (ipow(iabs(m) - iabs(m - k) - iabs(k)) * R[idx(n - j, m - k)]);
  };
  for (int j = 0; j <= P; ++j)
    for (int k = 0; k <= j; ++k)
      for (int n = j; n <= P; ++n)
        for (int m = 0; m <= n; ++m) pk.add(j, k, n, m, T(j, k, n, m), m > 0 ? T(j, k, n, -m) : qc{0, 0});
  pk.store(out);
}



// ---------------------------------------------------------------------------
// M2L by rotation: T(o) = H^T T(|o| z) D, with the rotation Q = R_y(-theta) R_z(-phi) taking the offset
// o = |o| (sin theta cos phi, sin theta sin phi, cos theta) onto +z. D (multipole coefficients) and H
// (L2P coefficient vectors) are the real representations of Q in the packed basis: packed-P2M(Q y) =
// D packed-P2M(y), L2P-row(Q x) = H L2P-row(x); since |Qx - Qy| = |x - y|, the local expansion of the
// rotated configuration L' = T(|o| z) D M satisfies L = H^T L'. Per degree n:
//   * R_z(-phi) rotates (Re, Im) of each order m by m phi: [[c, -s], [s, c]], c = cos m phi,
//     s = sin m phi, the same for D and H;
//   * R_y(-theta) commutes with the reflection y -> -y, so it maps cos-type coefficients
//     {Re (n, 0..n)} among themselves ((n+1)^2 block) and sin-type {Im (n, 1..n)} among themselves
//     (n^2 block);
//   * T(|o| z) couples (j, k) only with (n, k), the same block for the real and imaginary parts.
// The y-rotation blocks are found by probing the regular harmonics of degree n at 2n+1 pseudo-random
// points (the packed basis is the normalised one, so the systems are well conditioned) in FP64.
// Layout of `out` (m2l_rotation_size(P) doubles, blocks row-major [out][in]):
//   zc: (cos m phi, sin m phi) for m = 1..P;
//   for n = 0..P: Y_D cos block (n+1)^2, Y_D sin block n^2;
//   for m = 0..P: axial block (P+1-m)^2, indexed [j-m][n-m];
//   for n = 0..P: Y_H^T cos block (n+1)^2, Y_H^T sin block n^2.
// The map is then: D = Y_D Z, H^T = Z^T Y_H^T (Z from zc).
// ---------------------------------------------------------------------------
namespace {
void regular_d(double x, double y, double z, int L, std::vector<std::complex<double>>& R) {
  R.assign((L + 1) * (L + 1), 0.);
  std::vector<std::complex<double>> NR((L + 1) * (L + 1), 0.);
  const double r2 = x * x + y * y + z * z;
  NR[idx(0, 0)] = 1.;
  for (int m = 1; m <= L; ++m) NR[idx(m, m)] = (1. / (2 * m)) * (std::complex<double>(x, y) * NR[idx(m - 1, m - 1)]);
  for (int m = 0; m <= L; ++m)
    for (int n = m + 1; n <= L; ++n) {
      std::complex<double> a = (double)(2 * n - 1) * z * NR[idx(n - 1, m)];
      if (n - 2 >= m) a -= r2 * NR[idx(n - 2, m)];
      NR[idx(n, m)] = (1. / ((n + m) * (n - m))) * a;
    }
  for (int n = 0; n <= L; ++n)
    for (int m = 0; m <= n; ++m) {
      double c = (double)qt().sqfnm[n][m];
      if (m & 1) c = -c;
      R[idx(n, m)] = c * NR[idx(n, m)];
    }
}
// out = B A^-1 for d x d row-major A, B (partial pivoting)
void right_solve(std::vector<double> A, const std::vector<double>& B, int d, double* out) {
  std::vector<double> I(d * d, 0.);
  for (int i = 0; i < d; ++i) I[i * d + i] = 1.;
  for (int k = 0; k < d; ++k) {
    int p = k;
    for (int i = k + 1; i < d; ++i)
      if (std::fabs(A[i * d + k]) > std::fabs(A[p * d + k])) p = i;
    for (int j = 0; j < d; ++j) {
      std::swap(A[k * d + j], A[p * d + j]);
      std::swap(I[k * d + j], I[p * d + j]);
    }
    const double pv = A[k * d + k];
    for (int j = 0; j < d; ++j) {
      A[k * d + j] /= pv;
      I[k * d + j] /= pv;
    }
    for (int i = 0; i < d; ++i)
      if (i != k) {
        const double f = A[i * d + k];
        for (int j = 0; j < d; ++j) {
          A[i * d + j] -= f * A[k * d + j];
          I[i * d + j] -= f * I[k * d + j];
        }
      }
  }
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) {
      double s = 0.;
      for (int k = 0; k < d; ++k) s += B[i * d + k] * I[k * d + j];
      out[i * d + j] = s;
    }
}
// the (2n+1)^2 transform of degree n under x -> R_y(-theta) x: kind 0 = packed P2M vectors (Y_D),
// 1 = L2P rows (Y_H); row-major [out][in] over the packed degree block
void y_rotation(int n, double theta, int kind, std::vector<double>& X) {
  const int d = 2 * n + 1;
  std::uint64_t seed = 0x9e3779b97f4a7c15ull * (std::uint64_t)(2 * n + 1);
  auto unit = [&seed]() {
    std::uint64_t z = (seed += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return 2. * (double)(z >> 11) / 9007199254740992. - 1.;
  };
  const double c = std::cos(theta), s = std::sin(theta);
  std::vector<double> A(d * d), B(d * d);
  std::vector<std::complex<double>> Ra, Rb;
  for (int i = 0; i < d; ++i) {
    const double p[3] = {unit(), unit(), unit()};
    const double qp[3] = {c * p[0] - s * p[2], p[1], s * p[0] + c * p[2]};  // R_y(-theta) p
    regular_d(p[0], p[1], p[2], n, Ra);
    regular_d(qp[0], qp[1], qp[2], n, Rb);
    for (int m = 0; m <= n; ++m) {
      const std::complex<double> a = Ra[idx(n, m)], b = Rb[idx(n, m)];
      const double f = (kind == 1 && m > 0) ? 2. : 1.;  // L2P rows: 2 Re, -2 Im for m > 0
      const int re = m == 0 ? 0 : 2 * m - 1, im = 2 * m;
      A[re * d + i] = f * a.real();
      B[re * d + i] = f * b.real();
      if (m) {
        A[im * d + i] = -f * a.imag();
        B[im * d + i] = -f * b.imag();
      }
    }
  }
  X.resize(d * d);
  right_solve(A, B, d, X.data());
}
}  // namespace

// ---------------------------------------------------------------------------
// M2L through plane waves (exponential expansions; Greengard & Rokhlin 1997, section 7; the reference's
// lib_plane_wave_m2l.hpp:293-490, whose direction rotation is switched off at :306-313). For a box pair the
// translation t = target centre - source centre is turned by the rotation Q that takes its dominant axis
// (Up/Down/North/South/East/West) onto +z; in that frame, for z > 0 (our harmonics: Y_n^{-m} = conj(Y_n^m),
// I_n^m = Y_n^m / r^(n+1)),
//   I_n^m(x) = i^|m| / sqrt((n-|m|)!(n+|m|)!) int_0^inf lambda^n e^{-lambda z}
//              (1/2pi) int_0^2pi e^{i lambda (x cos a + y sin a)} e^{i m a} da dlambda,
//   e^{-lambda z + i lambda (x cos a + y sin a)} = sum_{n,m} (-lambda)^n i^|m| e^{-i m a} R_n^m(x) / sqrt((n-|m|)!(n+|m|)!),
// and the integrals are replaced by the generalized Gaussian rule in lambda (nodes, weights, M(k) angles of
// the published tables) and the trapezoidal rule in a. The multipole is converted to exponentials (M2X),
// translated diagonally (X2X: e^{-lambda t_z} e^{i lambda (t_x cos a + t_y sin a)}) and converted to a local
// expansion (X2L); the rules are designed for point separations 1 <= z <= 4, so lengths are divided by
// sigma = t'_z / 2.5 (multipole coefficients of degree n times sigma^-n, local ones sigma^-(n+1)). Here the
// whole chain of one lattice offset is composed into one packed C x C operator (rotation of the multipole,
// M2X, X2X, X2L, rotation back), evaluated in double: the GPU applies it as a dense contraction like the
// other M2L operators.
namespace {
struct PwNode {
  double lambda, weight;
  int M;
};
// Greengard & Rokhlin (1997), Tables 6 and 7: 6-digit (18 nodes, 558 exponentials) and 10-digit rules (30 nodes,
// 1751 exponentials); the same published values as lib_plane_wave_quadrature.hpp:41-100
const PwNode kPw6[18] = {
    {0.05278852766117, 0.13438265914335, 5},  {0.26949859838931, 0.29457752727395, 8},
    {0.63220353174689, 0.42607819361148, 12}, {1.11307564277608, 0.53189220776549, 16},
    {1.68939496140213, 0.61787306245538, 20}, {2.34376200469530, 0.68863156078905, 25},
    {3.06269982907806, 0.74749099381426, 29}, {3.83562941265296, 0.79699192718599, 34},
    {4.65424734321562, 0.83917454386997, 38}, {5.51209386593581, 0.87570092283745, 43},
    {6.40421268377278, 0.90792943590067, 47}, {7.32688001906175, 0.93698393742461, 51},
    {8.27740099258238, 0.96382546688788, 56}, {9.25397180602489, 0.98932985769673, 59},
    {10.25560272374640, 1.01438284597917, 59}, {11.28208829787774, 1.04003654374165, 51},
    {12.33406790967692, 1.06815489269567, 4}, {13.41492024017240, 1.10907580975537, 1}};
const PwNode kPw10[30] = {
    {0.03239542384523, 0.08289159611006, 7},   {0.16861844033714, 0.18838810673274, 10},
    {0.40611377169029, 0.28485143005306, 14},  {0.73466473057596, 0.37041553715895, 18},
    {1.14340561998398, 0.44539043894975, 22},  {1.62232408412252, 0.51100452150290, 26},
    {2.16276138867422, 0.56865283856139, 30},  {2.75739199003682, 0.61958013174010, 35},
    {3.40002470112078, 0.66481004321965, 39},  {4.08539104793552, 0.70517204769960, 43},
    {4.80897515497095, 0.74134967169016, 48},  {5.56688915983444, 0.77392103530415, 53},
    {6.35578243654166, 0.80338600122756, 57},  {7.17277232990713, 0.83018277269650, 62},
    {8.01538803542112, 0.85469824839953, 66},  {8.88152313049502, 0.87727539085565, 71},
    {9.76939480982937, 0.89821948245755, 76},  {10.67750922034750, 0.91780416582368, 80},
    {11.60463289992789, 0.93627766216629, 85}, {12.54977061299652, 0.95386940504388, 89},
    {13.51215012257297, 0.97079739700556, 94}, {14.49121482655196, 0.98727684670885, 97},
    {15.48662587630224, 1.00353112433459, 103}, {16.49827659770404, 1.01980697905712, 107},
    {17.52632405530625, 1.03639774457222, 110}, {18.57124579700721, 1.05368191266322, 112},
    {19.63393428118300, 1.07219343903929, 108}, {20.71585163675095, 1.09278318162014, 84},
    {21.81939113866225, 1.11737373706779, 4},  {22.95080495008893, 1.15786184931141, 1}};

// dense packed C x C matrices [in][out] of the rotation factors of m2l_rotation_operator(P, a): D (multipole
// into the frame where a is +z) and H^T (local back)
void rotation_dense(int P, int ax, int ay, int az, std::vector<double>& Dm, std::vector<double>& HTm) {
  const int C = (P + 1) * (P + 1);
  std::vector<double> f(m2l_rotation_size(P));
  m2l_rotation_operator(P, ax, ay, az, f.data());
  // Z (z-rotation per order, [[c, -s], [s, c]] on (Re, Im)) then Y (cos/sin blocks per degree)
  auto zrot = [&](std::vector<double>& v, bool transposed) {
    for (int n = 1; n <= P; ++n)
      for (int m = 1; m <= n; ++m) {
        const double c = f[2 * (m - 1)], sn = f[2 * (m - 1) + 1];
        const double re = v[pk_re(n, m)], im = v[pk_im(n, m)];
        if (!transposed) {
          v[pk_re(n, m)] = c * re - sn * im;
          v[pk_im(n, m)] = sn * re + c * im;
        } else {
          v[pk_re(n, m)] = c * re + sn * im;
          v[pk_im(n, m)] = -sn * re + c * im;
        }
      }
  };
  auto cs = [](int n, int kind, int a) { return kind == 0 ? pk_re(n, a) : pk_im(n, a + 1); };
  // offsets of the Y_D and Y_H^T blocks in f
  std::vector<int> yd(P + 1), yh(P + 1);
  int o = 2 * P;
  for (int n = 0; n <= P; ++n) {
    yd[n] = o;
    o += (n + 1) * (n + 1) + n * n;
  }
  for (int m = 0; m <= P; ++m) o += (P + 1 - m) * (P + 1 - m);
  for (int n = 0; n <= P; ++n) {
    yh[n] = o;
    o += (n + 1) * (n + 1) + n * n;
  }
  auto yblocks = [&](std::vector<double>& v, const std::vector<int>& base) {
    for (int n = 0; n <= P; ++n) {
      int b = base[n];
      for (int kind = 0; kind < 2; ++kind) {
        const int w = kind == 0 ? n + 1 : n;
        std::vector<double> out(w, 0.);
        for (int a = 0; a < w; ++a)
          for (int i = 0; i < w; ++i) out[a] += f[b + a * w + i] * v[cs(n, kind, i)];
        for (int a = 0; a < w; ++a) v[cs(n, kind, a)] = out[a];
        b += w * w;
      }
    }
  };
  Dm.assign((size_t)C * C, 0.);
  HTm.assign((size_t)C * C, 0.);
  for (int c = 0; c < C; ++c) {
    std::vector<double> v(C, 0.);
    v[c] = 1.;
    zrot(v, false);
    yblocks(v, yd);
    for (int r = 0; r < C; ++r) Dm[(size_t)c * C + r] = v[r];
    std::fill(v.begin(), v.end(), 0.);
    v[c] = 1.;
    yblocks(v, yh);
    zrot(v, true);
    for (int r = 0; r < C; ++r) HTm[(size_t)c * C + r] = v[r];
  }
}
}  // namespace

void m2l_plane_wave_operator(int P, int digits, int dx, int dy, int dz, double* out) {
  if (digits != 6 && digits != 10) throw std::runtime_error("plane-wave M2L: digits must be 6 or 10");
  const PwNode* rule = digits == 6 ? kPw6 : kPw10;
  const int nq = digits == 6 ? 18 : 30;
  const int C = (P + 1) * (P + 1);
  // t = target - source = -o; the axis of its largest component, signed, is turned onto +z
  const double t[3] = {-(double)dx, -(double)dy, -(double)dz};
  int ax = 2;
  if (std::fabs(t[0]) > std::fabs(t[2]) && std::fabs(t[0]) >= std::fabs(t[1])) ax = 0;
  else if (std::fabs(t[1]) > std::fabs(t[2])) ax = 1;
  int a[3] = {0, 0, 0};
  a[ax] = t[ax] > 0 ? 1 : -1;
  // m2l_rotation_operator(P, a) rotates the offset a onto +z: Q = R_y(-theta) R_z(-phi)
  const double theta = std::acos((double)a[2]), phi = std::atan2((double)a[1], (double)a[0]);
  const double cp = std::cos(phi), sp = std::sin(phi), ct = std::cos(theta), st = std::sin(theta);
  const double u0 = cp * t[0] + sp * t[1], u1 = -sp * t[0] + cp * t[1], u2 = t[2];  // R_z(-phi) t
  const double tr[3] = {ct * u0 - st * u2, u1, st * u0 + ct * u2};                   // R_y(-theta)
  const double sigma = tr[2] / 2.5;
  const double ts[3] = {tr[0] / sigma, tr[1] / sigma, tr[2] / sigma};
  // F_q(d) = sum over the angles of node q of e^{i d a} e^{i lambda (ts_x cos a + ts_y sin a)}, d = m - k
  const int D = 2 * P;
  std::vector<std::complex<double>> F((size_t)nq * (2 * D + 1));
  for (int q = 0; q < nq; ++q) {
    std::complex<double>* Fq = F.data() + (size_t)q * (2 * D + 1);
    for (int j = 0; j < rule[q].M; ++j) {
      const double al = 2. * M_PI * j / rule[q].M;
      const std::complex<double> base =
          std::exp(std::complex<double>(0., rule[q].lambda * (ts[0] * std::cos(al) + ts[1] * std::sin(al))));
      const std::complex<double> rot(std::cos(al), std::sin(al));
      std::complex<double> up = base, dn = base;  // base e^{i d al} for d = 0, 1, .. and 0, -1, ..
      Fq[D] += base;
      for (int d = 1; d <= D; ++d) {
        up *= rot;
        dn *= std::conj(rot);
        Fq[D + d] += up;
        Fq[D - d] += dn;
      }
    }
    const double f = (rule[q].weight / rule[q].M) * std::exp(-rule[q].lambda * ts[2]);
    for (int d = 0; d <= 2 * D; ++d) Fq[d] *= f;
  }
  auto ipw = [](int e) {  // i^e
    const int r = ((e % 4) + 4) % 4;
    return r == 0 ? std::complex<double>(1, 0) : r == 1 ? std::complex<double>(0, 1) : r == 2 ? std::complex<double>(-1, 0) : std::complex<double>(0, -1);
  };
  auto nrm = [](int n, int m) { return 1. / std::sqrt(std::tgamma(n - std::abs(m) + 1.) * std::tgamma(n + std::abs(m) + 1.)); };
  // lambda_q^n (-lambda_q)^j sigma^-(n+j+1) per node and degree pair
  std::vector<double> lp((size_t)nq * (P + 1) * (P + 1));
  for (int q = 0; q < nq; ++q)
    for (int n = 0; n <= P; ++n)
      for (int j = 0; j <= P; ++j)
        lp[((size_t)q * (P + 1) + n) * (P + 1) + j] =
            std::pow(rule[q].lambda, n) * std::pow(-rule[q].lambda, j) * std::pow(sigma, -(n + j + 1));
  // T[(j,k),(n,m)] in the rotated, scaled frame
  Packer pk(P);
  auto T = [&](int j, int k, int n, int m) -> qc {
    std::complex<double> acc = 0.;
    for (int q = 0; q < nq; ++q)
      acc += lp[((size_t)q * (P + 1) + n) * (P + 1) + j] * F[(size_t)q * (2 * D + 1) + (m - k) + D];
    const std::complex<double> v = acc * ipw(std::abs(m) + std::abs(k)) * (nrm(n, m) * nrm(j, k));
    return qc{(q)v.real(), (q)v.imag()};
  };
  for (int j = 0; j <= P; ++j)
    for (int k = 0; k <= j; ++k)
      for (int n = 0; n <= P; ++n)
        for (int m = 0; m <= n; ++m) pk.add(j, k, n, m, T(j, k, n, m), m > 0 ? T(j, k, n, -m) : qc{0, 0});
  std::vector<double> X((size_t)C * C);
  pk.store(X.data());
  // out = H^T X D (all [in][out]): out[c][r] = sum_a sum_b D[c][a] X[a][b] H^T[b][r]
  std::vector<double> Dm, HTm, XD((size_t)C * C, 0.);
  rotation_dense(P, a[0], a[1], a[2], Dm, HTm);
  for (int c = 0; c < C; ++c)
    for (int aa = 0; aa < C; ++aa) {
      const double d = Dm[(size_t)c * C + aa];
      if (d == 0.) continue;
      for (int b = 0; b < C; ++b) XD[(size_t)c * C + b] += d * X[(size_t)aa * C + b];
    }
  // H^T maps within a degree: the packed entries of degree n are n^2 .. (n+1)^2 - 1
  for (int c = 0; c < C; ++c)
    for (int r = 0; r < C; ++r) {
      const int n = (int)std::sqrt((double)r + 0.5);
      double acc = 0.;
      for (int b = n * n; b < (n + 1) * (n + 1); ++b) acc += XD[(size_t)c * C + b] * HTm[(size_t)b * C + r];
      out[(size_t)c * C + r] = acc;
    }
}

int m2l_rotation_size(int P) {
  int s = 2 * P;
  for (int n = 0; n <= P; ++n) s += 2 * ((n + 1) * (n + 1) + n * n);
  for (int m = 0; m <= P; ++m) s += (P + 1 - m) * (P + 1 - m);
  return s;
}

void m2l_rotation_operator(int P, int dx, int dy, int dz, double* out) {
  const double r = std::sqrt((double)dx * dx + (double)dy * dy + (double)dz * dz);
  const double theta = std::acos(std::max(-1., std::min(1., dz / r))), phi = std::atan2((double)dy, (double)dx);
  int o = 0;
  for (int m = 1; m <= P; ++m) {
    out[o++] = std::cos(m * phi);
    out[o++] = std::sin(m * phi);
  }
  // split a packed degree block into its cos-type (Re of m = 0..n) and sin-type (Im of m = 1..n) blocks
  auto split = [&](int n, const std::vector<double>& X, bool transpose) {
    const int d = 2 * n + 1;
    auto at = [&](int a, int b) { return transpose ? X[b * d + a] : X[a * d + b]; };
    for (int a = 0; a <= n; ++a)
      for (int b = 0; b <= n; ++b) out[o++] = at(a == 0 ? 0 : 2 * a - 1, b == 0 ? 0 : 2 * b - 1);
    for (int a = 1; a <= n; ++a)
      for (int b = 1; b <= n; ++b) out[o++] = at(2 * a, 2 * b);
  };
  std::vector<double> Y;
  for (int n = 0; n <= P; ++n) {
    y_rotation(n, theta, 0, Y);
    split(n, Y, false);
  }
  // axial operator T(r z): per order m, block [j - m][n - m] (j, n = m..P), shared by Re and Im parts
  const std::vector<qc> I = irregular(0, 0, (q)r, 2 * P);
  for (int m = 0; m <= P; ++m)
    for (int j = m; j <= P; ++j)
      for (int n = m; n <= P; ++n) {
        const q a = Anm(n, m) * Anm(j, m) / (((n & 1) ? (q)-1 : (q)1) * Anm(j + n, 0));
        out[o++] = (double)(a * (ipow(-2 * m) * I[idx(j + n, 0)])).re;
      }
  for (int n = 0; n <= P; ++n) {
    y_rotation(n, theta, 1, Y);
    split(n, Y, true);
  }
}

}  // namespace bemfmm
