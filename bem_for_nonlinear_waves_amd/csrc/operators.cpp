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

#include <cmath>
#include <complex>
#include <cstdint>
#include <vector>

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
// M2L by rotation: T(o) = H^T T(|o| z) D (checked in float128 to 1e-34 relative).
// D (multipole coefficients) and H (L2P coefficient vectors) are the block-diagonal (one (2n+1)^2
// block per degree) real representations of the rotation Q with Q o = |o| z: packed-P2M(Q y) =
// D packed-P2M(y), L2P-row(Q x) = H L2P-row(x); since |Qx - Qy| = |x - y|, the local expansion of
// the rotated configuration L' = T(|o| z) D M satisfies L = H^T L'. T(|o| z) couples (j,k) only with
// (n,k) (harmonics on the z axis have order 0), the same block for the real and imaginary parts.
// The blocks are found by probing: the regular harmonics of degree n at 2n+1 points (the basis is the
// normalised one, so the probe systems are well conditioned) - double precision is enough for the
// FP32 operator these factors serve.
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
}  // namespace

int m2l_rotation_size(int P) {
  int s = 0;
  for (int n = 0; n <= P; ++n) s += 2 * (2 * n + 1) * (2 * n + 1);
  for (int m = 0; m <= P; ++m) s += (P + 1 - m) * (P + 1 - m);
  return s;
}

void m2l_rotation_operator(int P, int dx, int dy, int dz, double* out) {
  const double r = std::sqrt((double)dx * dx + (double)dy * dy + (double)dz * dz);
  const double ez[3] = {dx / r, dy / r, dz / r};
  double t[3] = {1., 0., 0.};
  if (std::fabs(ez[0]) > 0.9) {
    t[0] = 0.;
    t[1] = 1.;
  }
  const double dt = t[0] * ez[0] + t[1] * ez[1] + t[2] * ez[2];
  double ex[3] = {t[0] - dt * ez[0], t[1] - dt * ez[1], t[2] - dt * ez[2]};
  const double nx = std::sqrt(ex[0] * ex[0] + ex[1] * ex[1] + ex[2] * ex[2]);
  for (double& v : ex) v /= nx;
  const double ey[3] = {ez[1] * ex[2] - ez[2] * ex[1], ez[2] * ex[0] - ez[0] * ex[2], ez[0] * ex[1] - ez[1] * ex[0]};
  // probe points: fixed pseudo-random points of the cube [-1, 1]^3 (splitmix64)
  std::uint64_t seed = 0x9e3779b97f4a7c15ull;
  auto unit = [&seed]() {
    std::uint64_t z = (seed += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return 2. * (double)(z >> 11) / 9007199254740992. - 1.;
  };
  auto probe = [&unit](int, int, double* p) {
    for (int a = 0; a < 3; ++a) p[a] = unit();
  };
  double* D = out;
  int off = 0;
  for (int n = 0; n <= P; ++n) off += (2 * n + 1) * (2 * n + 1);
  double* Tax = out + off;
  for (int m = 0; m <= P; ++m) off += (P + 1 - m) * (P + 1 - m);
  double* Ht = out + off;
  std::vector<std::complex<double>> Ra, Rb;
  int boff = 0;
  for (int n = 0; n <= P; ++n) {
    const int d = 2 * n + 1;
    std::vector<double> Am(d * d), Bm(d * d), Al(d * d), Bl(d * d);
    for (int i = 0; i < d; ++i) {
      double pnt[3], qp[3];
      probe(i, d, pnt);
      for (int a = 0; a < 3; ++a) qp[a] = 0.;
      for (int a = 0; a < 3; ++a) {
        qp[0] += ex[a] * pnt[a];
        qp[1] += ey[a] * pnt[a];
        qp[2] += ez[a] * pnt[a];
      }
      regular_d(pnt[0], pnt[1], pnt[2], n, Ra);
      regular_d(qp[0], qp[1], qp[2], n, Rb);
      for (int m = 0; m <= n; ++m) {
        const std::complex<double> a = Ra[idx(n, m)], b = Rb[idx(n, m)];
        const int re = m == 0 ? 0 : 2 * m - 1, im = 2 * m;  // offsets of pk_re / pk_im inside the degree block
        Am[re * d + i] = a.real();
        Bm[re * d + i] = b.real();
        Al[re * d + i] = (m ? 2. : 1.) * a.real();
        Bl[re * d + i] = (m ? 2. : 1.) * b.real();
        if (m) {
          Am[im * d + i] = -a.imag();
          Bm[im * d + i] = -b.imag();
          Al[im * d + i] = -2. * a.imag();
          Bl[im * d + i] = -2. * b.imag();
        }
      }
    }
    std::vector<double> Hn(d * d);
    right_solve(Am, Bm, d, D + boff);  // D_n (row-major [out][in])
    right_solve(Al, Bl, d, Hn.data());
    for (int a = 0; a < d; ++a)
      for (int c = 0; c < d; ++c) Ht[boff + a * d + c] = Hn[c * d + a];  // H_n^T (row-major [out][in])
    boff += d * d;
  }
  // axial operator T(r z): per order m, block [j - m][n - m] (j, n = m..P), shared by Re and Im parts
  const std::vector<qc> I = irregular(0, 0, (q)r, 2 * P);
  int toff = 0;
  for (int m = 0; m <= P; ++m) {
    const int w = P + 1 - m;
    for (int j = m; j <= P; ++j)
      for (int n = m; n <= P; ++n) {
        const q a = Anm(n, m) * Anm(j, m) / (((n & 1) ? (q)-1 : (q)1) * Anm(j + n, 0));
        const qc v = a * (ipow(-2 * m) * I[idx(j + n, 0)]);
        Tax[toff + (j - m) * w + (n - m)] = (double)v.re;
      }
    toff += w * w;
  }
}

}  // namespace bemfmm
