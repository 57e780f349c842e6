// ORACLE — test infrastructure only. NOT part of the product.
//
// FP64 CPU restatement of the reference's FMM-accelerated GMRES BVP path
// (tomoakihirakawa/BEM_for_Nonlinear_Waves), used by tests/, bench.py's
// cpu_baseline leg and __graft_entry__.smoke() as the checker. The product
// (libbemfmm.so) never links, loads or calls this code.
//
// Every stage follows the reference file:line cited at its definition:
//   scaling / orchestration   bem/BEM_solveBVP_GMRES_FMM.hpp:3056-3506
//   sources (linear faces)    bem/BEM_solveBVP_GMRES_FMM.hpp:1799-1937
//   octree (Buckets)          lib/include/lib_spatial_partitioning.hpp:223-1307
//   MAC + lists               lib/include/moments.hpp:948-997
//   P2M/M2M/M2L/L2L/L2P       lib/include/moments.hpp:165-885, lib_multipole_expansion.hpp:779-1150
//   near field                lib/include/lib_multipole_expansion.hpp:1180-1302
//   matvec                    bem/BEM_solveBVP_GMRES_FMM.hpp:34-158
//   rigid-mode alpha          bem/BEM_solveBVP_GMRES_FMM.hpp:2674-2716
//   GMRES                     bem/BEM_solveBVP_GMRES_FMM.hpp:2007-2355, lib/include/basic_linear_systems.hpp:3137-3451
//   ILU(0) preconditioner     bem/BEM_solveBVP_GMRES_FMM.hpp:215-662, lib/include/basic_linear_systems.hpp:2337-2603
//   SCHWARZ preconditioner    bem/BEM_solveBVP_GMRES_FMM.hpp:901-1013, lib/include/basic_linear_systems.hpp:2800-3003
//
// The reference's translation tables AAA_{M2M,M2L,L2L}_FMM.hpp are absent, so
// the three operators are evaluated from the Greengard & Rokhlin (1997)
// closed forms (Theorems 3.3, 3.5, 3.6) with the coefficient A_n^m in long
// double; the reference's harmonics use libstdc++ std::sph_legendre, which is
// what this file uses too.
//
// Pinning status of each function is listed in the file header of
// tests/test_oracle.py (known-answer tests, closed forms, invariants).

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <deque>
#include <map>
#include <set>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using cd = std::complex<double>;
using V3 = std::array<double, 3>;

inline V3 operator-(const V3& a, const V3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline V3 operator+(const V3& a, const V3& b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
inline V3 operator*(double s, const V3& a) { return {s * a[0], s * a[1], s * a[2]}; }
inline double dot(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline double norm(const V3& a) { return std::hypot(a[0], a[1], a[2]); }  // Norm: basic_arithmetic_array_operations.hpp:860
inline V3 cross(const V3& a, const V3& b) { return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]}; }

// ---------------------------------------------------------------------------
// Quadrature tables (values as printed in lib/include/basic_constants.hpp:1326-1350
// and lib/include/dunavant_rules.hpp:18-46; checked against Gauss-Legendre and
// Dunavant monomial exactness in tests/test_oracle.py).
// ---------------------------------------------------------------------------
const double kGW5[25][3] = {
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

const std::vector<std::array<double, 4>>& dunavant(int q) {
  static const std::vector<std::array<double, 4>> d1 = {{1. / 3., 1. / 3., 1. / 3., 0.5}};
  static const std::vector<std::array<double, 4>> d3 = {
      {2. / 3., 1. / 6., 1. / 6., 1. / 6.}, {1. / 6., 2. / 3., 1. / 6., 1. / 6.}, {1. / 6., 1. / 6., 2. / 3., 1. / 6.}};
  static const std::vector<std::array<double, 4>> d6 = {
      {0.816847572980459, 0.091576213509771, 0.091576213509771, 0.054975871827661},
      {0.091576213509771, 0.816847572980459, 0.091576213509771, 0.054975871827661},
      {0.091576213509771, 0.091576213509771, 0.816847572980459, 0.054975871827661},
      {0.108103018168070, 0.445948490915965, 0.445948490915965, 0.111690794839006},
      {0.445948490915965, 0.108103018168070, 0.445948490915965, 0.111690794839006},
      {0.445948490915965, 0.445948490915965, 0.108103018168070, 0.111690794839006}};
  static const std::vector<std::array<double, 4>> d7 = {
      {1. / 3., 1. / 3., 1. / 3., 0.1125},
      {0.059715871789770, 0.470142064105115, 0.470142064105115, 0.066197076394253},
      {0.470142064105115, 0.059715871789770, 0.470142064105115, 0.066197076394253},
      {0.470142064105115, 0.470142064105115, 0.059715871789770, 0.066197076394253},
      {0.797426985353087, 0.101286507323456, 0.101286507323456, 0.062969590272414},
      {0.101286507323456, 0.797426985353087, 0.101286507323456, 0.062969590272414},
      {0.101286507323456, 0.101286507323456, 0.797426985353087, 0.062969590272414}};
  switch (q) {
  case 1: return d1;
  case 3: return d3;
  case 6: return d6;
  case 7: return d7;
  default: throw std::runtime_error("oracle: unsupported p2m_quadrature_points");
  }
}

// ---------------------------------------------------------------------------
// Spherical harmonics, reference convention (lib_multipole_expansion.hpp:578-614, 878-890):
//   Y_n^m = sqrt(4 pi/(2n+1)) * std::sph_legendre(n,|m|,theta) * exp(i m phi)
//   R_n^m(x) = rho^n Y_n^m,   I_n^m(x) = Y_n^m / rho^(n+1)
// Coordinates (lib_multipole_expansion.hpp:788-799): rho = hypot(x,y,z),
// theta = atan2(hypot(x,y), z), phi = atan2(y, x).
// Storage: idx(n,m) = n*(n+1)+m for -n<=m<=n.
// ---------------------------------------------------------------------------
inline int idx(int n, int m) { return n * (n + 1) + m; }

struct Harm {
  int P;
  std::vector<cd> Y;  // Y_n^m for n<=P
  double rho;
  void eval(const V3& x, int P_) {
    P = P_;
    Y.assign((P + 1) * (P + 1), 0.);
    rho = norm(x);
    const double r2d = std::hypot(x[0], x[1]);
    const double theta = std::atan2(r2d, x[2]);
    const double phi = std::atan2(x[1], x[0]);
    const double two_pi_sqrt = std::sqrt(2.0 * M_PI);
    for (int n = 0; n <= P; ++n) {
      const double s = two_pi_sqrt / std::sqrt(n + 0.5);
      for (int m = 0; m <= n; ++m) {
        const double a = std::sph_legendre(n, m, theta) * s;
        Y[idx(n, m)] = std::polar(a, m * phi);
        Y[idx(n, -m)] = std::polar(a, -m * phi);
      }
    }
  }
};

// Regular solid harmonics R_n^m and their normal derivatives grad(R_n^{-m}).nhat,
// as used by P2M (lib_multipole_expansion.hpp:996-1027, m -> -m).  The gradient is
// evaluated with the ladder identities of the unnormalised harmonic
// NR_n^m = rho^n P_n^m(cos t) e^{i m phi}/(n+m)!:
//   d/dz NR_n^m = NR_{n-1}^m, (dx + i dy) NR_n^m = -NR_{n-1}^{m+1}, (dx - i dy) NR_n^m = NR_{n-1}^{m-1},
// and R_n^q = c(n,q) NR_n^q with c(n,q) = (-1)^max(q,0) sqrt((n-|q|)!(n+|q|)!).
// This is the same derivative the reference writes with the polar-angle chain
// rule, but stays finite on the z axis.
double lfact(int k) { return std::lgamma(k + 1.0); }
struct CfacTable {
  double v[64][64];
  CfacTable() {
    for (int n = 0; n < 64; ++n)
      for (int a = 0; a <= n; ++a) v[n][a] = std::exp(0.5 * (lfact(n - a) + lfact(n + a)));
  }
};
double cfac(int n, int q) {
  static const CfacTable t;
  const int a = std::abs(q);
  const double v = t.v[n][a];
  return (q > 0 && (q & 1)) ? -v : v;
}

struct SolidR {
  int P;
  std::vector<cd> R;   // R_n^m(x), all m
  std::vector<cd> NR;  // NR_n^m
  void eval(const V3& x, int P_) {
    P = P_;
    Harm h;
    h.eval(x, P);
    R.assign((P + 1) * (P + 1), 0.);
    NR.assign((P + 1) * (P + 1), 0.);
    for (int n = 0; n <= P; ++n) {
      const double rn = std::pow(h.rho, n);
      for (int m = -n; m <= n; ++m) {
        R[idx(n, m)] = h.Y[idx(n, m)] * rn;
        NR[idx(n, m)] = R[idx(n, m)] / cfac(n, m);
      }
    }
  }
  cd nr(int n, int m) const { return (n < 0 || std::abs(m) > n) ? cd(0.) : NR[idx(n, m)]; }
  // grad(R_n^q) . nhat
  cd dR(int n, int q, const V3& nh) const {
    if (n == 0) return 0.;
    const cd a(nh[0] * 0.5, -nh[1] * 0.5);  // (nx - i ny)/2
    const cd b(nh[0] * 0.5, nh[1] * 0.5);   // (nx + i ny)/2
    const cd g = -a * nr(n - 1, q + 1) + b * nr(n - 1, q - 1) + nh[2] * nr(n - 1, q);
    return cfac(n, q) * g;
  }
};

// ---------------------------------------------------------------------------
// Greengard & Rokhlin (1997) translation coefficients; A_n^m = (-1)^n / sqrt((n-m)!(n+m)!)
// (lib_multipole_expansion_constants.hpp:14-28).  Index maps follow moments.hpp:204-270.
// ---------------------------------------------------------------------------
long double Anm(int n, int m) {
  const int a = std::abs(m);
  if (a > n) return 0.L;
  long double f1 = 1.L, f2 = 1.L;
  for (int k = 2; k <= n - a; ++k) f1 *= k;
  for (int k = 2; k <= n + a; ++k) f2 *= k;
  const long double v = 1.L / std::sqrt(f1 * f2);
  return (n & 1) ? -v : v;
}
cd ipow(int e) {
  e = ((e % 4) + 4) % 4;
  static const cd t[4] = {cd(1, 0), cd(0, 1), cd(-1, 0), cd(0, -1)};
  return t[e];
}
// M2M: M_j^k(parent) += sum_{n,m} AAA_M2M[j][k][n][m] * R_n^{-m}(child-parent) * M_{j-n}^{k-m}(child)
cd aaa_m2m(int j, int k, int n, int m) {
  if (std::abs(k - m) > j - n || std::abs(m) > n) return 0.;
  const long double v = Anm(n, m) * Anm(j - n, k - m) / Anm(j, k);
  return ipow(std::abs(k) - std::abs(m) - std::abs(k - m)) * (double)v;
}
// M2L: L_j^k(target) += sum_{n,m} AAA_M2L[j][k][n][m] * I_{j+n}^{m-k}(source-target) * M_n^m(source)
cd aaa_m2l(int j, int k, int n, int m) {
  const long double v = Anm(n, m) * Anm(j, k) / (((n & 1) ? -1.L : 1.L) * Anm(j + n, m - k));
  return ipow(std::abs(k - m) - std::abs(k) - std::abs(m)) * (double)v;
}
// L2L: L_j^k(child) += sum_{n>=j,m} AAA_L2L[j][k][n][m] * R_{n-j}^{m-k}(parent-child) * L_n^m(parent)
cd aaa_l2l(int j, int k, int n, int m) {
  if (std::abs(m - k) > n - j) return 0.;
  const long double v = Anm(n - j, m - k) * Anm(j, k) / ((((n + j) & 1) ? -1.L : 1.L) * Anm(n, m));
  return ipow(std::abs(m) - std::abs(m - k) - std::abs(k)) * (double)v;
}

struct Tables {
  int P;
  std::vector<cd> m2m, m2l, l2l;  // [j][k][n][m] flattened, (P+1)^2 x (P+1)^2
  int C;
  void build(int P_) {
    P = P_;
    C = (P + 1) * (P + 1);
    m2m.assign((size_t)C * C, 0.);
    m2l.assign((size_t)C * C, 0.);
    l2l.assign((size_t)C * C, 0.);
    for (int j = 0; j <= P; ++j)
      for (int k = -j; k <= j; ++k)
        for (int n = 0; n <= P; ++n)
          for (int m = -n; m <= n; ++m) {
            const size_t o = (size_t)idx(j, k) * C + idx(n, m);
            if (n <= j) m2m[o] = aaa_m2m(j, k, n, m);
            m2l[o] = aaa_m2l(j, k, n, m);
            if (n >= j) l2l[o] = aaa_l2l(j, k, n, m);
          }
  }
};

// ---------------------------------------------------------------------------
// Octree (Buckets), lib/include/lib_spatial_partitioning.hpp:223-1307.
// Each node is a grid of ceil(width/dL) cells per axis (2 in practice); a child's
// bounds are its parent's cell and its cell size is parent.dL*(0.5+1e-13)
// (lib_spatial_partitioning.hpp:451-456, 1087-1094).
// ---------------------------------------------------------------------------
struct Bucket {
  int level = 0, max_level = 0;
  double bmin[3], bmax[3];
  V3 X;  // centre = mean of bounds (basic_geometry.hpp:633-700)
  double dL = 0, inv_dL = 0;
  int size[3] = {1, 1, 1};
  Bucket* parent = nullptr;
  std::vector<Bucket*> cells;  // size[0]*size[1]*size[2], x-major
  std::vector<std::vector<int>> cell_data;
  std::vector<int> data;  // data1D_vector: all descendant faces (insertion order)
  double off = 0.;        // max_internal_offset
  std::vector<Bucket*> m2l_list, near, l2m;
  std::vector<int> l2m_op;  // cached-M2L mode: operator index per l2m entry
  std::vector<cd> M0, M1, L0, L1;
  int id = -1;
  bool has_children() const {
    for (auto* c : cells)
      if (c) return true;
    return false;
  }
  static int calc_size(double lo, double hi, double dl) {
    double s = std::ceil((hi - lo) / dl);
    if (s <= 0.) s = 1.;
    return (int)s;
  }
  void init(const double lo[3], const double hi[3], double dl) {
    for (int a = 0; a < 3; ++a) {
      bmin[a] = lo[a];
      bmax[a] = hi[a];
      X[a] = (lo[a] + hi[a]) / 2.;
      size[a] = calc_size(lo[a], hi[a], dl);
    }
    dL = dl;
    inv_dL = 1.0 / dl;
    cell_data.assign(size[0] * size[1] * size[2], {});
    cells.assign(size[0] * size[1] * size[2], nullptr);
  }
  void index_nc(const V3& x, int ijk[3]) const {
    for (int a = 0; a < 3; ++a) ijk[a] = (int)std::floor((x[a] - bmin[a]) * inv_dL);
  }
  bool inside(const V3& x) const {
    int ijk[3];
    index_nc(x, ijk);
    for (int a = 0; a < 3; ++a)
      if (ijk[a] < 0 || ijk[a] >= size[a]) return false;
    return true;
  }
  int cell_of(const V3& x) const {
    int ijk[3];
    index_nc(x, ijk);
    for (int a = 0; a < 3; ++a) ijk[a] = std::clamp(ijk[a], 0, size[a] - 1);
    return (ijk[0] * size[1] + ijk[1]) * size[2] + ijk[2];
  }
  void cell_bounds(int c, double lo[3], double hi[3]) const {
    int k = c % size[2], j = (c / size[2]) % size[1], i = c / (size[2] * size[1]);
    const int ijk[3] = {i, j, k};
    for (int a = 0; a < 3; ++a) {
      lo[a] = dL * ijk[a] + bmin[a];
      hi[a] = dL * (ijk[a] + 1) + bmin[a];
    }
  }
};

struct FaceRec {
  int npf = 3;  // 3: linear triangle; 6: true quadratic face (vertices, then the midpoints of edges 01, 12, 20)
  int v[6];
  int dof[6];   // -1: edge midpoint without a DOF (its weight goes half to each end vertex)
  V3 X[6];
  V3 centroid;  // (a+b+c)/3
  V3 pole;      // Dot({1/3,1/3,1/3}, X012): tree position
  V3 nhat;
  V3 cr;        // 2*area*normal
  double J;     // 2*area
  double maxoff;
  std::vector<V3> qx;
  std::vector<double> qw;                  // W = J * w_q (J at the source point for quadratic faces)
  std::vector<std::array<double, 3>> qb;   // barycentric
  std::vector<std::array<double, 6>> qN;   // density weights per face node (the barycentrics, or TriShape<6>)
  std::vector<V3> qn;                      // unit normal at the source point
};

// ---- true quadratic faces (BEM_solveBVP_GMRES_FMM.hpp:1311-1657) ----
// TriShape<6> and its derivatives, t2 = 1 - t0 - t1 (basic_arithmetic_array_operations.hpp:1467-1566)
void q6_shape(double t0, double t1, double N[6], double D0[6], double D1[6]) {
  const double t2 = 1. - t0 - t1;
  const double n[6] = {t0 * (2. * t0 - 1.), t1 * (2. * t1 - 1.), t2 * (2. * t2 - 1.), 4. * t0 * t1, 4. * t1 * t2,
                       4. * t2 * t0};
  const double a[6] = {-1. + 4. * t0, 0., 1. - 4. * t2, 4. * t1, -4. * t1, -4. * t0 + 4. * t2};
  const double b[6] = {0., -1. + 4. * t1, 1. - 4. * t2, 4. * t0, 4. * t2 - 4. * t1, -4. * t0};
  for (int k = 0; k < 6; ++k) {
    N[k] = n[k];
    D0[k] = a[k];
    D1[k] = b[k];
  }
}
// X(t0, t1) and dX/dt0 x dX/dt1 of the 6-node geometry
V3 q6_point(const FaceRec& R, double t0, double t1, V3& crs) {
  double N[6], D0[6], D1[6];
  q6_shape(t0, t1, N, D0, D1);
  V3 x{0., 0., 0.}, a{0., 0., 0.}, b{0., 0., 0.};
  for (int k = 0; k < 6; ++k) {
    x = x + N[k] * R.X[k];
    a = a + D0[k] * R.X[k];
    b = b + D1[k] * R.X[k];
  }
  crs = cross(a, b);
  return x;
}
// trueQuadN6 (Network.hpp:1607-1632)
std::array<double, 6> q6_density(const FaceRec& R, double t0, double t1) {
  double N[6], D0[6], D1[6];
  q6_shape(t0, t1, N, D0, D1);
  std::array<double, 6> out{N[0], N[1], N[2], N[3], N[4], N[5]};
  for (int e = 0; e < 3; ++e)
    if (R.dof[3 + e] < 0) {
      out[e] += 0.5 * out[3 + e];
      out[(e + 1) % 3] += 0.5 * out[3 + e];
      out[3 + e] = 0.;
    }
  return out;
}

struct Oracle {
  // inputs
  int n_nodes = 0, n_faces = 0, n_dof = 0;
  std::vector<V3> Xn;  // scaled node coordinates
  std::vector<FaceRec> F;
  std::vector<int> dof_kind, dof_point;
  std::vector<std::vector<int>> point_dofs;
  std::vector<double> phi_known, phin_known;  // scaled knowns
  double Lscale = 1.0, scale = 1.0, near_region = 0.;
  // settings
  int P = 10, leaf_max = 50, max_level = 7, q = 6, m2l_mode = 0;
  int npf = 3;  // nodes per face (6: true quadratic)
  double theta = 0.25;
  // tree
  Bucket root;
  std::vector<std::unique_ptr<Bucket>> pool;
  std::vector<std::vector<Bucket*>> level_buckets;
  std::vector<Bucket*> leaves;
  std::vector<Bucket*> target_box;  // per node
  Tables T;
  // near field per target node
  std::vector<std::vector<int>> near_idx;
  std::vector<std::vector<double>> near_wG, near_wGn;
  std::vector<double> alpha;  // per DOF
  // working dense arrays
  std::vector<double> phi, phin;
  // stats
  long long n_m2l_pairs = 0;
  long long nnz = 0;
  double t_setup = 0;

  // ---- tree ----------------------------------------------------------------
  Bucket* new_bucket() {
    pool.emplace_back(new Bucket());
    return pool.back().get();
  }
  void add(Bucket* b, int f) {
    const V3& x = F[f].pole;
    b->data.push_back(f);
    b->cell_data[b->cell_of(x)].push_back(f);
  }
  // generateTree (lib_spatial_partitioning.hpp:1079-1115) with the grow condition
  // of BEM_solveBVP_GMRES_FMM.hpp:1199.
  void grow(Bucket* b) {
    const bool cond = !b->data.empty() && (int)b->data.size() > leaf_max && b->level < b->max_level;
    if (!(cond && b->level < b->max_level)) return;
    for (size_t c = 0; c < b->cells.size(); ++c) {
      if (b->cell_data[c].empty()) continue;
      double lo[3], hi[3];
      b->cell_bounds((int)c, lo, hi);
      Bucket* ch = new_bucket();
      ch->init(lo, hi, b->dL * (0.5 + 1e-13));
      ch->parent = b;
      ch->level = b->level + 1;
      ch->max_level = b->max_level;
      for (int f : b->cell_data[c])
        if (ch->inside(F[f].pole)) add(ch, f);
      b->cells[c] = ch;
      grow(ch);
    }
  }
  void collect(Bucket* b) {
    // rebuildHierarchyLists (lib_spatial_partitioning.hpp:966-984): preorder traversal
    for (auto* c : b->cells) {
      if (!c) continue;
      if (!c->data.empty()) {
        level_buckets[c->level].push_back(c);
        if (!c->has_children()) leaves.push_back(c);
      }
      collect(c);
    }
  }
  Bucket* deepest(const V3& x) {
    // getBucketAtDeepest (lib_spatial_partitioning.hpp:1270-1279)
    Bucket* cur = &root;
    while (cur->has_children()) {
      Bucket* nx = cur->cells[cur->cell_of(x)];
      if (!nx) return cur;
      cur = nx;
    }
    return cur;
  }
  bool is_far(const Bucket* A, const Bucket* B) const {
    // moments.hpp:948-954
    const double eff = std::max(A->dL + 2.0 * A->off, B->dL + 2.0 * B->off);
    return theta > std::sqrt(3.0) * eff / norm(B->X - A->X);
  }
  void lists(Bucket* A, Bucket* B) {
    // buildNearAndM2LLists (moments.hpp:968-979)
    if (!is_far(A, B)) {
      if (A != B) A->near.push_back(B);
      for (auto* Ac : A->cells) {
        if (!Ac) continue;
        for (auto* Bc : B->cells) {
          if (!Bc) continue;
          if (Ac != Bc && is_far(Ac, Bc))
            Ac->m2l_list.push_back(Bc);
          else
            lists(Ac, Bc);
        }
      }
    }
  }

  // ---- setup ----------------------------------------------------------------
  void setup(const double* X, const int* tri, const int* fdof, const int* kind, const int* dpoint, const double* phi_in,
             const double* phin_in) {
    // Coordinate scaling: L = bounding-box diagonal (Network.hpp:3215-3232; GMRES_FMM.hpp:3076-3091)
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < n_nodes; ++i)
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], X[3 * i + a]);
        hi[a] = std::max(hi[a], X[3 * i + a]);
      }
    Lscale = std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) + (hi[2] - lo[2]) * (hi[2] - lo[2]));
    Xn.resize(n_nodes);
    double slo[3] = {1e300, 1e300, 1e300}, shi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < n_nodes; ++i)
      for (int a = 0; a < 3; ++a) {
        Xn[i][a] = X[3 * i + a] / Lscale;
        slo[a] = std::min(slo[a], Xn[i][a]);
        shi[a] = std::max(shi[a], Xn[i][a]);
      }
    scale = norm(V3{shi[0] - slo[0], shi[1] - slo[1], shi[2] - slo[2]});  // getScale (basic_geometry.hpp:736-739)
    near_region = scale / 25.;
    // DOFs, knowns (Neumann phin scaled by L: copyToFMM, GMRES_FMM.hpp:1203-1250)
    dof_kind.assign(kind, kind + n_dof);
    dof_point.assign(dpoint, dpoint + n_dof);
    point_dofs.assign(n_nodes, {});
    for (int i = 0; i < n_dof; ++i) point_dofs[dof_point[i]].push_back(i);
    phi_known.assign(phi_in, phi_in + n_dof);
    phin_known.resize(n_dof);
    for (int i = 0; i < n_dof; ++i) phin_known[i] = (kind[i] != 0) ? phin_in[i] * Lscale : phin_in[i];
    // Faces and P2M sources (GMRES_FMM.hpp:1799-1937)
    const auto& rule = dunavant(q);
    F.resize(n_faces);
    for (int f = 0; f < n_faces; ++f) {
      FaceRec& R = F[f];
      R.npf = npf;
      for (int k = 0; k < npf; ++k) {
        R.v[k] = tri[npf * f + k];
        R.dof[k] = fdof[npf * f + k];
        R.X[k] = Xn[R.v[k]];
      }
      const V3 c = cross(R.X[1] - R.X[0], R.X[2] - R.X[0]);
      R.J = norm(c);
      R.nhat = (1.0 / R.J) * c;
      R.cr = R.J * R.nhat;
      for (int a = 0; a < 3; ++a) {
        R.centroid[a] = (R.X[0][a] + R.X[1][a] + R.X[2][a]) / 3.;
        R.pole[a] = (1. / 3.) * R.X[0][a] + (1. / 3.) * R.X[1][a] + (1. / 3.) * R.X[2][a];
      }
      if (npf == 6) R.pole = R.centroid;  // pole.X = F->centroid for quadratic faces
      R.qx.clear();
      R.qw.clear();
      R.qb.clear();
      R.qN.clear();
      R.qn.clear();
      R.maxoff = 0.;
      for (const auto& r : rule) {
        V3 xq;
        if (npf == 3) {
          for (int a = 0; a < 3; ++a) xq[a] = r[0] * R.X[0][a] + r[1] * R.X[1][a] + r[2] * R.X[2][a];
          R.qw.push_back(R.J * r[3]);
          R.qN.push_back({r[0], r[1], r[2], 0., 0., 0.});
          R.qn.push_back(R.nhat);
        } else {
          V3 crs;
          xq = q6_point(R, r[0], r[1], crs);
          const double Jq = norm(crs);
          R.qw.push_back(Jq * r[3]);
          R.qN.push_back(q6_density(R, r[0], r[1]));
          R.qn.push_back(Jq > 1e-30 ? (1.0 / Jq) * crs : R.nhat);
        }
        R.qx.push_back(xq);
        R.qb.push_back({r[0], r[1], r[2]});
        R.maxoff = std::max(R.maxoff, norm(xq - R.pole));
      }
    }
    // Root bucket (initializeBucket, GMRES_FMM.hpp:1178-1201)
    {
      V3 Cn;
      for (int a = 0; a < 3; ++a) Cn[a] = (slo[a] + shi[a]) / 2.;
      const double dL = 1.05 * scale / 2.;
      double blo[3], bhi[3];
      for (int a = 0; a < 3; ++a) {
        blo[a] = Cn[a] - dL;
        bhi[a] = Cn[a] + dL;
      }
      root.init(blo, bhi, dL);
      root.level = 0;
      root.max_level = max_level;
    }
    for (int f = 0; f < n_faces; ++f)
      if (root.inside(F[f].pole)) add(&root, f);
    grow(&root);
    level_buckets.assign(max_level + 1, {});
    leaves.clear();
    if (!root.data.empty()) {
      level_buckets[0].push_back(&root);
      if (!root.has_children()) leaves.push_back(&root);
    }
    collect(&root);
    int id = 0;
    for (auto& lv : level_buckets)
      for (auto* b : lv) b->id = id++;
    // max_internal_offset (moments.hpp:1323-1337)
    for (auto* b : leaves) {
      b->off = 0.;
      for (int f : b->data) b->off = std::max(b->off, F[f].maxoff);
    }
    for (int lv = max_level - 1; lv >= 0; --lv)
      for (auto* b : level_buckets[lv]) {
        if (!b->has_children()) continue;
        b->off = 0.;
        for (auto* c : b->cells)
          if (c) b->off = std::max(b->off, c->off);
      }
    // lists (setBucketsForM2L, moments.hpp:983-997)
    for (auto* A : root.cells) {
      if (!A) continue;
      for (auto* B : root.cells)
        if (B) lists(A, B);
    }
    for (auto& lv : level_buckets)
      for (auto* A : lv)
        for (auto* B : A->m2l_list) B->l2m.push_back(A);
    n_m2l_pairs = 0;
    for (auto& lv : level_buckets)
      for (auto* A : lv) n_m2l_pairs += (long long)A->l2m.size();
    const int C = (P + 1) * (P + 1);
    for (auto& b : pool) {
      b->M0.assign(C, 0.);
      b->M1.assign(C, 0.);
      b->L0.assign(C, 0.);
      b->L1.assign(C, 0.);
    }
    root.M0.assign(C, 0.);
    root.M1.assign(C, 0.);
    root.L0.assign(C, 0.);
    root.L1.assign(C, 0.);
    T.build(P);
    // targets
    target_box.resize(n_nodes);
    for (int i = 0; i < n_nodes; ++i) target_box[i] = deepest(Xn[i]);
    build_near();
    phi.assign(n_dof, 0.);
    phin.assign(n_dof, 0.);
    compute_alpha();
  }

  // ---- near field (setDirectIntegration, lib_multipole_expansion.hpp:1180-1271) ----
  void face_direct(int f, int t, std::vector<double>& accG, std::vector<double>& accGn, std::vector<int>& touched,
                   std::vector<char>& flag) const {
    const FaceRec& R = F[f];
    const V3& x = Xn[t];
    auto addw = [&](int d, double g, double gn) {
      accG[d] += g;
      accGn[d] += gn;
      if (!flag[d]) {
        flag[d] = 1;
        touched.push_back(d);
      }
    };
    if (R.npf == 6) {
      face_direct_quad(R, t, addw);
      return;
    }
    int slot = -1;
    for (int k = 0; k < 3; ++k)
      if (R.v[k] == t) slot = k;
    double WG[3] = {0, 0, 0}, WGn[3] = {0, 0, 0};
    if (slot >= 0) {
      // adjacent: fill_direct_linear (GMRES_FMM.hpp:1824-1860), GW5x5 collapsed at the target vertex
      int perm[3];  // rotated vertex r -> original corner
      if (slot == 0) { perm[0] = 0; perm[1] = 1; perm[2] = 2; }
      else if (slot == 1) { perm[0] = 1; perm[1] = 2; perm[2] = 0; }
      else { perm[0] = 2; perm[1] = 0; perm[2] = 1; }
      double G[3] = {0, 0, 0}, Gn[3] = {0, 0, 0};
      for (const auto& w : kGW5) {
        const double t0 = w[0], t1 = w[1], ww = w[2];
        const double N[3] = {t0, t1 * (1. - t0), (t0 - 1.) * (t1 - 1.)};
        V3 xq;
        for (int a = 0; a < 3; ++a) xq[a] = N[0] * R.X[perm[0]][a] + N[1] * R.X[perm[1]][a] + N[2] * R.X[perm[2]][a];
        const V3 Rv = xq - x;
        const double nr_inv = 1. / norm(Rv);
        const double tmp = ww * (1. - t0) * nr_inv;
        const double rc = -dot(nr_inv * Rv, R.cr);
        for (int r = 0; r < 3; ++r) {
          G[r] += R.J * (tmp * N[r]);
          Gn[r] += rc * (tmp * N[r]) * nr_inv;
        }
      }
      for (int r = 0; r < 3; ++r) {
        WG[perm[r]] = G[r];
        WGn[perm[r]] = Gn[r];
      }
      WGn[0] = 0.0;  // acc.add(idx0, WGN[idx[0]], 0.0): corner-0 Gn dropped
    } else {
      // non-adjacent: fill_nonadj_linear (GMRES_FMM.hpp:1863-1891)
      if (norm(R.centroid - x) < near_region) {
        for (const auto& w : kGW5) {
          const double t0 = w[0], t1 = w[1], ww = w[2];
          const double N[3] = {t0, t1 * (1. - t0), (t0 - 1.) * (t1 - 1.)};
          V3 xq;
          for (int a = 0; a < 3; ++a) xq[a] = N[0] * R.X[0][a] + N[1] * R.X[1][a] + N[2] * R.X[2][a];
          const V3 Rv = xq - x;
          const double nr_inv = 1. / norm(Rv);
          const double rc = dot(nr_inv * Rv, R.cr);
          for (int r = 0; r < 3; ++r) {
            const double tmp = (ww * (1. - t0) * N[r]) * nr_inv;
            WG[r] += R.J * tmp;
            WGn[r] -= rc * (tmp * nr_inv);
          }
        }
      } else {
        for (size_t s = 0; s < R.qx.size(); ++s) {
          const V3 Rv = R.qx[s] - x;
          const double nr_inv = 1. / norm(Rv);
          const double rc = dot(nr_inv * Rv, R.cr);
          const double ww = R.qw[s] / R.J;
          for (int r = 0; r < 3; ++r) {
            const double tmp = ww * R.qb[s][r] * nr_inv;
            WG[r] += R.J * tmp;
            WGn[r] -= rc * (tmp * nr_inv);
          }
        }
      }
    }
    for (int r = 0; r < 3; ++r) addw(R.dof[r], WG[r], WGn[r]);
  }

  // fill_direct_true_quad (vertex and edge-midpoint Duffy, GMRES_FMM.hpp:1424-1558) and fill_nonadj_true_quad
  // (:1562-1592): weights per face node, the target node's own double-layer weight dropped
  template <class Add>
  void face_direct_quad(const FaceRec& R, int t, Add&& addw) const {
    const V3& x = Xn[t];
    int slot = -1;
    for (int k = 0; k < 6; ++k)
      if (R.v[k] == t) slot = k;
    double WG[6] = {0, 0, 0, 0, 0, 0}, WGn[6] = {0, 0, 0, 0, 0, 0};
    auto at = [&](double b0, double b1, double w, int drop) {
      V3 crs;
      const V3 xq = q6_point(R, b0, b1, crs);
      const auto N = q6_density(R, b0, b1);
      const V3 Rv = xq - x;
      const double r = norm(Rv);
      const double g = norm(crs) * w / r;
      const double gn = -dot((1.0 / r) * Rv, crs) * w / (r * r);
      for (int k = 0; k < 6; ++k) {
        WG[k] += g * N[k];
        if (k != drop) WGn[k] += gn * N[k];
      }
    };
    if (slot >= 0 && slot < 3) {
      for (const auto& w : kGW5) {
        const double t0 = w[0], t1 = w[1], ww = w[2];
        double b[3];
        b[slot] = t0;
        b[(slot + 1) % 3] = t1 * (1. - t0);
        b[(slot + 2) % 3] = (1. - t0) * (1. - t1);
        at(b[0], b[1], ww * (1. - t0), slot);
      }
    } else if (slot >= 3) {
      const int e = slot - 3, va = e, vb = (e + 1) % 3, vc = (e + 2) % 3;
      for (const auto& w : kGW5) {
        const double t0 = w[0], t1 = w[1], ww = w[2];
        const double o = 1. - t0, s1 = t1 * o, s2 = o * (1. - t1);
        double b[3];
        b[va] = 0.5 * t0;
        b[vb] = 0.5 * t0 + s1;
        b[vc] = s2;
        at(b[0], b[1], ww * o * 0.5, slot);
        b[va] = 0.5 * t0 + s2;
        b[vb] = 0.5 * t0;
        b[vc] = s1;
        at(b[0], b[1], ww * o * 0.5, slot);
      }
    } else {
      const auto& rule = norm(R.centroid - x) < near_region ? dunavant(q) : dunavant(6);
      for (const auto& r : rule) at(r[0], r[1], r[3], -1);
    }
    for (int k = 0; k < 6; ++k)
      if (R.dof[k] >= 0) addw(R.dof[k], WG[k], WGn[k]);
  }

  std::vector<int> near_faces_of(const Bucket* bd) const {
    std::vector<int> out;
    auto setNear = [&](const Bucket* B) { out.insert(out.end(), B->data.begin(), B->data.end()); };
    setNear(bd);
    for (auto* b : bd->near) setNear(b);
    for (const Bucket* b = bd->parent; b; b = b->parent)
      for (auto* B : b->near)
        if (!B->has_children()) setNear(B);
    return out;
  }

  void build_near() {
    near_idx.assign(n_nodes, {});
    near_wG.assign(n_nodes, {});
    near_wGn.assign(n_nodes, {});
#pragma omp parallel
    {
      std::vector<double> accG(n_dof, 0.), accGn(n_dof, 0.);
      std::vector<char> flag(n_dof, 0);
      std::vector<int> touched;
#pragma omp for schedule(dynamic, 16)
      for (int t = 0; t < n_nodes; ++t) {
        touched.clear();
        for (int f : near_faces_of(target_box[t])) face_direct(f, t, accG, accGn, touched, flag);
        std::sort(touched.begin(), touched.end());
        for (int d : touched) {
          near_idx[t].push_back(d);
          near_wG[t].push_back(accG[d]);
          near_wGn[t].push_back(accGn[d]);
          accG[d] = accGn[d] = 0.;
          flag[d] = 0;
        }
      }
    }
    nnz = 0;
    for (auto& v : near_idx) nnz += (long long)v.size();
  }

  // ---- FMM passes (updateFMM, moments.hpp:1547-1574) ---------------------------
  void p2m() {
    const int C = (P + 1) * (P + 1);
#pragma omp parallel for schedule(dynamic, 4)
    for (size_t li = 0; li < leaves.size(); ++li) {
      Bucket* b = leaves[li];
      std::fill(b->M0.begin(), b->M0.end(), cd(0.));
      std::fill(b->M1.begin(), b->M1.end(), cd(0.));
      SolidR S;
      for (int f : b->data) {
        const FaceRec& R = F[f];
        for (size_t s = 0; s < R.qx.size(); ++s) {
          double sphi = 0., sphin = 0.;
          for (int k = 0; k < R.npf; ++k)
            if (R.dof[k] >= 0) {
              sphi += R.qN[s][k] * phi[R.dof[k]];
              sphin += R.qN[s][k] * phin[R.dof[k]];
            }
          const double w_phi = R.qw[s] * sphi, w_phin = R.qw[s] * sphin;
          S.eval(R.qx[s] - b->X, P);
          for (int n = 0; n <= P; ++n)
            for (int m = 0; m <= n; ++m) {
              b->M0[idx(n, m)] += w_phin * S.R[idx(n, -m)];
              b->M1[idx(n, m)] += w_phi * S.dR(n, -m, R.qn[s]);
            }
        }
      }
      conj_fill(b->M0);
      conj_fill(b->M1);
    }
    (void)C;
  }
  void conj_fill(std::vector<cd>& v) const {
    for (int n = 0; n <= P; ++n)
      for (int m = 1; m <= n; ++m) v[idx(n, -m)] = std::conj(v[idx(n, m)]);
  }
  void m2m() {
    // M2M per level bottom-up (moments.hpp:401-436, 912-915); real-field rows k>=0.
    const int C = (P + 1) * (P + 1);
    for (int lv = max_level - 1; lv >= 0; --lv) {
#pragma omp parallel for schedule(dynamic, 4)
      for (size_t bi = 0; bi < level_buckets[lv].size(); ++bi) {
        Bucket* b = level_buckets[lv][bi];
        if (!b->has_children()) continue;
        std::fill(b->M0.begin(), b->M0.end(), cd(0.));
        std::fill(b->M1.begin(), b->M1.end(), cd(0.));
        SolidR S;
        for (auto* c : b->cells) {
          if (!c) continue;
          S.eval(c->X - b->X, P);
          for (int j = 0; j <= P; ++j)
            for (int k = 0; k <= j; ++k) {
              cd s0 = 0., s1 = 0.;
              for (int n = 0; n <= j; ++n)
                for (int m = -n; m <= n; ++m) {
                  if (std::abs(k - m) > j - n) continue;
                  const cd a = T.m2m[(size_t)idx(j, k) * C + idx(n, m)] * S.R[idx(n, -m)];
                  s0 += a * c->M0[idx(j - n, k - m)];
                  s1 += a * c->M1[idx(j - n, k - m)];
                }
              b->M0[idx(j, k)] += s0;
              b->M1[idx(j, k)] += s1;
            }
        }
        conj_fill(b->M0);
        conj_fill(b->M1);
      }
    }
  }
  // Cached-M2L mode (CPU baseline at large N): the per-pair coefficients a = AAA_M2L * I_{j+n}^{m-k}(d)
  // that SimpleM2L stores per pair (moments.hpp:493-598) are computed once per (level, lattice offset)
  // from a representative pair and reused by every pair with that offset; the sums are unchanged.
  bool m2l_cached = false;
  std::vector<std::vector<cd>> m2l_ops;  // [(j,k>=0) row][(n,m) col]
  void m2l_coefficients(const V3& d, std::vector<cd>& out) const {
    const int C = (P + 1) * (P + 1);
    Harm H;
    std::vector<cd> Iv((2 * P + 1) * (2 * P + 1));
    H.eval(d, 2 * P);
    for (int l = 0; l <= 2 * P; ++l) {
      const double rl = 1.0 / (std::pow(H.rho, l) * H.rho);
      for (int qq = -l; qq <= l; ++qq) Iv[idx(l, qq)] = H.Y[idx(l, qq)] * rl;
    }
    out.assign((size_t)(P + 1) * (P + 2) / 2 * C, cd(0.));
    int row = 0;
    for (int j = 0; j <= P; ++j)
      for (int k = 0; k <= j; ++k, ++row)
        for (int n = 0; n <= P; ++n)
          for (int m = -n; m <= n; ++m)
            out[(size_t)row * C + idx(n, m)] = T.m2l[(size_t)idx(j, k) * C + idx(n, m)] * Iv[idx(j + n, m - k)];
  }
  void build_m2l_cache(const std::vector<Bucket*>& all) {
    std::map<std::array<long long, 4>, int> key;
    std::vector<V3> rep;
    for (Bucket* A : all) {
      A->l2m_op.clear();
      const double side = A->bmax[0] - A->bmin[0];
      for (Bucket* B : A->l2m) {
        const V3 d = B->X - A->X;
        const std::array<long long, 4> k{A->level, std::llround(d[0] / side), std::llround(d[1] / side),
                                         std::llround(d[2] / side)};
        auto it = key.find(k);
        if (it == key.end()) {
          it = key.emplace(k, (int)rep.size()).first;
          rep.push_back(d);
        }
        A->l2m_op.push_back(it->second);
      }
    }
    m2l_ops.assign(rep.size(), {});
#pragma omp parallel for schedule(dynamic, 1)
    for (size_t i = 0; i < rep.size(); ++i) m2l_coefficients(rep[i], m2l_ops[i]);
  }
  void m2l() {
    // SimpleM2L (moments.hpp:493-598, 732-764): L_j^k += sum AAA_M2L * I_{j+n}^{m-k}(s-t) * M_n^m(s)
    const int C = (P + 1) * (P + 1);
    std::vector<Bucket*> all;
    for (auto& lv : level_buckets)
      for (auto* b : lv) all.push_back(b);
    if (m2l_cached) {
      if (m2l_ops.empty()) build_m2l_cache(all);
#pragma omp parallel for schedule(dynamic, 1)
      for (size_t ai = 0; ai < all.size(); ++ai) {
        Bucket* A = all[ai];
        std::fill(A->L0.begin(), A->L0.end(), cd(0.));
        std::fill(A->L1.begin(), A->L1.end(), cd(0.));
        for (size_t q = 0; q < A->l2m.size(); ++q) {
          const Bucket* B = A->l2m[q];
          const cd* a = m2l_ops[A->l2m_op[q]].data();
          int row = 0;
          for (int j = 0; j <= P; ++j)
            for (int k = 0; k <= j; ++k, ++row) {
              cd s0 = 0., s1 = 0.;
              const cd* ar = a + (size_t)row * C;
              for (int c = 0; c < C; ++c) {
                s0 += ar[c] * B->M0[c];
                s1 += ar[c] * B->M1[c];
              }
              A->L0[idx(j, k)] += s0;
              A->L1[idx(j, k)] += s1;
            }
        }
        conj_fill(A->L0);
        conj_fill(A->L1);
      }
      return;
    }
    // per-(exact offset) operator cache for large problems (m2l_mode 1)
#pragma omp parallel for schedule(dynamic, 1)
    for (size_t ai = 0; ai < all.size(); ++ai) {
      Bucket* A = all[ai];
      std::fill(A->L0.begin(), A->L0.end(), cd(0.));
      std::fill(A->L1.begin(), A->L1.end(), cd(0.));
      Harm H;
      std::vector<cd> Iv((2 * P + 1) * (2 * P + 1));
      for (auto* B : A->l2m) {
        const V3 d = B->X - A->X;
        H.eval(d, 2 * P);
        for (int l = 0; l <= 2 * P; ++l) {
          const double rl = 1.0 / (std::pow(H.rho, l) * H.rho);
          for (int qq = -l; qq <= l; ++qq) Iv[idx(l, qq)] = H.Y[idx(l, qq)] * rl;
        }
        for (int j = 0; j <= P; ++j)
          for (int k = 0; k <= j; ++k) {
            cd s0 = 0., s1 = 0.;
            for (int n = 0; n <= P; ++n)
              for (int m = -n; m <= n; ++m) {
                const cd a = T.m2l[(size_t)idx(j, k) * C + idx(n, m)] * Iv[idx(j + n, m - k)];
                s0 += a * B->M0[idx(n, m)];
                s1 += a * B->M1[idx(n, m)];
              }
            A->L0[idx(j, k)] += s0;
            A->L1[idx(j, k)] += s1;
          }
      }
      conj_fill(A->L0);
      conj_fill(A->L1);
    }
  }
  void l2l() {
    // L2L top-down (moments.hpp:862-884, 925-929)
    const int C = (P + 1) * (P + 1);
    for (int lv = 0; lv <= max_level; ++lv) {
#pragma omp parallel for schedule(dynamic, 4)
      for (size_t bi = 0; bi < level_buckets[lv].size(); ++bi) {
        Bucket* p = level_buckets[lv][bi];
        SolidR S;
        for (auto* c : p->cells) {
          if (!c) continue;
          S.eval(p->X - c->X, P);
          for (int j = 0; j <= P; ++j)
            for (int k = 0; k <= j; ++k) {
              cd s0 = 0., s1 = 0.;
              for (int n = j; n <= P; ++n)
                for (int m = -n; m <= n; ++m) {
                  if (std::abs(m - k) > n - j) continue;
                  const cd a = T.l2l[(size_t)idx(j, k) * C + idx(n, m)] * S.R[idx(n - j, m - k)];
                  s0 += a * p->L0[idx(n, m)];
                  s1 += a * p->L1[idx(n, m)];
                }
              c->L0[idx(j, k)] += s0;
              c->L1[idx(j, k)] += s1;
            }
          conj_fill(c->L0);
          conj_fill(c->L1);
        }
      }
    }
  }
  void update_fmm() {
    p2m();
    m2m();
    m2l();
    l2l();
  }
  // near + far per target node (compute_Ax_minus_b inner loop, GMRES_FMM.hpp:62-108)
  void eval_targets(std::vector<double>& gphin, std::vector<double>& gnphi) {
    gphin.assign(n_nodes, 0.);
    gnphi.assign(n_nodes, 0.);
#pragma omp parallel for schedule(static)
    for (int t = 0; t < n_nodes; ++t) {
      double n0 = 0., n1 = 0.;
      for (size_t i = 0; i < near_idx[t].size(); ++i) {
        n0 += phin[near_idx[t][i]] * near_wG[t][i];
        n1 += phi[near_idx[t][i]] * near_wGn[t][i];
      }
      const Bucket* b = target_box[t];
      SolidR S;
      S.eval(Xn[t] - b->X, P);
      cd f0 = 0., f1 = 0.;
      for (int n = 0; n <= P; ++n)
        for (int m = -n; m <= n; ++m) {
          f0 += S.R[idx(n, m)] * b->L0[idx(n, m)];
          f1 += S.R[idx(n, m)] * b->L1[idx(n, m)];
        }
      gphin[t] = n0 + f0.real();
      gnphi[t] = n1 + f1.real();
    }
  }
  void set_unknowns(const double* x) {
    for (int i = 0; i < n_dof; ++i) {
      if (dof_kind[i] == 0) phin[i] = x[i];
      else phi[i] = x[i];
    }
  }
  void set_knowns_zero() {
    for (int i = 0; i < n_dof; ++i) {
      if (dof_kind[i] == 0) phi[i] = 0.;
      else phin[i] = 0.;
    }
  }
  void set_knowns() {
    for (int i = 0; i < n_dof; ++i) {
      if (dof_kind[i] == 0) phi[i] = phi_known[i];
      else phin[i] = phin_known[i];
    }
  }
  // ROBIN DOFs (kind 3): phi_n = kappa phi + g. The postprocess of the frequency-domain solver sets
  // phin = kappa * phi on the is_robin DOFs after the unknowns are placed (main_freq_domain.cpp:543-555,
  // g = 0 there); the known part g enters with the other knowns. kappa_s = kappa L (scaled phi_n).
  std::vector<double> kappa_s;
  void add_robin() {
    if (kappa_s.empty()) return;
    for (int i = 0; i < n_dof; ++i)
      if (dof_kind[i] == 3) phin[i] += kappa_s[i] * phi[i];
  }
  void set_robin(const double* kappa) {
    kappa_s.assign(n_dof, 0.);
    for (int i = 0; i < n_dof; ++i)
      if (dof_kind[i] == 3) kappa_s[i] = kappa[i] * Lscale;
  }
  double kap(int j) const { return kappa_s.empty() ? 0. : kappa_s[j]; }
  void compute_alpha() {
    // rigid mode (GMRES_FMM.hpp:2674-2716): phi=1, phin=0 on all DOFs; alpha = -GnPhi
    alpha.assign(n_dof, 1.0);
    std::fill(phi.begin(), phi.end(), 1.0);
    std::fill(phin.begin(), phin.end(), 0.0);
    update_fmm();
    std::vector<double> g0, g1;
    eval_targets(g0, g1);
    for (int i = 0; i < n_dof; ++i) {
      double a = -g1[dof_point[i]];
      if (std::abs(a) < 1e-15) a = 1.0;
      alpha[i] = a;
    }
  }
  // compute_Ax_minus_b (GMRES_FMM.hpp:34-158) with the current phi/phin
  void ax_minus_b(std::vector<double>& Av) {
    update_fmm();
    std::vector<double> g0, g1;
    eval_targets(g0, g1);
    Av.assign(n_dof, 0.);
    for (int i = 0; i < n_dof; ++i) {
      const int t = dof_point[i];
      if (dof_kind[i] == 2) Av[i] = phi[i];
      else Av[i] = g0[t] - (g1[t] + alpha[i] * phi[i]);
    }
  }
  void matvec(const double* x, double* y) {
    set_unknowns(x);
    set_knowns_zero();
    add_robin();
    std::vector<double> Av;
    ax_minus_b(Av);
    std::copy(Av.begin(), Av.end(), y);
  }
  void rhs(double* b) {
    // solveSystemGMRES: setUnknowns(0); b = -compute_Ax_minus_b(); CORNER rows <- phi_D (GMRES_FMM.hpp:2025-2047)
    std::vector<double> z(n_dof, 0.);
    set_unknowns(z.data());
    set_knowns();
    std::vector<double> Av;
    ax_minus_b(Av);
    for (int i = 0; i < n_dof; ++i) b[i] = (dof_kind[i] == 2) ? phi_known[i] : -Av[i];
  }

  // ---- FMM-consistent dense operator: near weights + direct sum over far P2M sources
  void dense(double* A) {
    std::fill(A, A + (size_t)n_dof * n_dof, 0.);
    std::vector<double> rowG(n_dof), rowGn(n_dof);
    for (int t = 0; t < n_nodes; ++t) {
      std::fill(rowG.begin(), rowG.end(), 0.);
      std::fill(rowGn.begin(), rowGn.end(), 0.);
      for (size_t i = 0; i < near_idx[t].size(); ++i) {
        rowG[near_idx[t][i]] += near_wG[t][i];
        rowGn[near_idx[t][i]] += near_wGn[t][i];
      }
      std::vector<char> isnear(n_faces, 0);
      for (int f : near_faces_of(target_box[t])) isnear[f] = 1;
      for (int f = 0; f < n_faces; ++f) {
        if (isnear[f]) continue;
        const FaceRec& R = F[f];
        for (size_t s = 0; s < R.qx.size(); ++s) {
          const V3 Rv = R.qx[s] - Xn[t];
          const double r = norm(Rv);
          const double g = R.qw[s] / r;
          const double gn = -R.qw[s] * dot(Rv, R.qn[s]) / (r * r * r);
          for (int k = 0; k < R.npf; ++k) {
            if (R.dof[k] < 0) continue;
            rowG[R.dof[k]] += g * R.qN[s][k];
            rowGn[R.dof[k]] += gn * R.qN[s][k];
          }
        }
      }
      for (int i : point_dofs[t]) {
        double* row = A + (size_t)i * n_dof;
        if (dof_kind[i] == 2) {
          row[i] = 1.0;
          continue;
        }
        for (int j = 0; j < n_dof; ++j) {
          // unknown of DOF j: phin (Dirichlet) enters through G, phi (Neumann) through -Gn, phi (ROBIN)
          // through kappa_j G - Gn
          row[j] = (dof_kind[j] == 0) ? rowG[j] : -rowGn[j];
          if (dof_kind[j] == 3) row[j] += kap(j) * rowG[j];
        }
        // alpha*phi_i is an unknown term only for rows whose phi is unknown (Neumann, ROBIN)
        if (dof_kind[i] == 1 || dof_kind[i] == 3) row[i] -= alpha[i];
      }
    }
  }

  // ---- GMRES: run_gmres control flow (GMRES_FMM.hpp:2275-2355) with the gmres
  // class (basic_linear_systems.hpp:3137-3451): MGS Arnoldi, Givens QR.
  // Diagonal of the operator (Dirichlet: G_ii, Neumann: -Gn_ii - alpha_i, CORNER: 1) for left Jacobi scaling.
  std::vector<double> diag_inv() const {
    std::vector<double> d(n_dof, 1.0);
    for (int i = 0; i < n_dof; ++i) {
      if (dof_kind[i] == 2) continue;
      const int t = dof_point[i];
      const auto& ix = near_idx[t];
      const size_t k = std::lower_bound(ix.begin(), ix.end(), i) - ix.begin();
      double g = 1.0;
      if (k < ix.size() && ix[k] == i) {
        g = dof_kind[i] == 0 ? near_wG[t][k] : -near_wGn[t][k] - alpha[i];
        if (dof_kind[i] == 3) g += kap(i) * near_wG[t][k];
      }
      d[i] = (std::isfinite(g) && std::abs(g) > 1e-15) ? 1.0 / g : 1.0;
    }
    return d;
  }
  // ---- restricted additive Schwarz over given overlapping blocks (SchwarzAdditive_CRS,
  // basic_linear_systems.hpp:2870-3003, with the sparse matrix of GMRES_FMM.hpp:440-640).
  // A_b = operator restricted to block b with the near-field entries of each row (entries of
  // DOFs outside the row point's near set are 0), -alpha_i on Neumann diagonals, CORNER rows e_i.
  // Each block is solved by partial-pivot LU; z takes the block's solution on the block's core.
  std::vector<int64_t> sw_ptr{0};
  std::vector<int32_t> sw_idx, sw_ncore;
  double near_entry(int i, int j) const {
    if (dof_kind[i] == 2) return i == j ? 1.0 : 0.0;
    const int t = dof_point[i];
    const auto& ix = near_idx[t];
    const size_t k = std::lower_bound(ix.begin(), ix.end(), j) - ix.begin();
    double v = 0.;
    if (k < ix.size() && ix[k] == j) {
      v = dof_kind[j] == 0 ? near_wG[t][k] : -near_wGn[t][k];
      if (dof_kind[j] == 3) v += kap(j) * near_wG[t][k];
    }
    if (i == j && (dof_kind[i] == 1 || dof_kind[i] == 3)) v -= alpha[i];
    return v;
  }
  void schwarz_apply(const double* r, double* z) const {
    for (int i = 0; i < n_dof; ++i) z[i] = 0.;
    const int nb = (int)sw_ncore.size();
    for (int b = 0; b < nb; ++b) {
      const int64_t p0 = sw_ptr[b];
      const int m = (int)(sw_ptr[b + 1] - p0);
      const int32_t* id = &sw_idx[p0];
      std::vector<double> A((size_t)m * m), y(m);
      for (int i = 0; i < m; ++i) {
        y[i] = r[id[i]];
        for (int j = 0; j < m; ++j) A[(size_t)i * m + j] = near_entry(id[i], id[j]);
      }
      // LU with partial pivoting, forward/back substitution
      for (int k = 0; k < m; ++k) {
        int pr = k;
        for (int i = k + 1; i < m; ++i)
          if (std::abs(A[(size_t)i * m + k]) > std::abs(A[(size_t)pr * m + k])) pr = i;
        if (pr != k) {
          for (int j = 0; j < m; ++j) std::swap(A[(size_t)k * m + j], A[(size_t)pr * m + j]);
          std::swap(y[k], y[pr]);
        }
        const double pv = A[(size_t)k * m + k];
        for (int i = k + 1; i < m; ++i) {
          const double f = A[(size_t)i * m + k] / pv;
          A[(size_t)i * m + k] = f;
          for (int j = k + 1; j < m; ++j) A[(size_t)i * m + j] -= f * A[(size_t)k * m + j];
          y[i] -= f * y[k];
        }
      }
      for (int i = m - 1; i >= 0; --i) {
        double s = y[i];
        for (int j = i + 1; j < m; ++j) s -= A[(size_t)i * m + j] * y[j];
        y[i] = s / A[(size_t)i * m + i];
      }
      for (int i = 0; i < sw_ncore[b]; ++i) z[id[i]] = y[i];
    }
  }

  // ---- the reference's sparse preconditioner matrix and its ILU(0) --------------------------------------
  // buildSparseMatrixForILU (bem/BEM_solveBVP_GMRES_FMM.hpp:215-662, linear elements): per boundary point, the
  // faces of its neighbourhood, each integrated with networkFace::setIntegrationInfo's rules
  // (lib/src/networkFace.cpp:28-85) seen from the face's points in getPoints(origin) order
  // (lib/include/Network.hpp:2062-2069); ILU0_CRS (lib/include/basic_linear_systems.hpp:2337-2603).
  std::vector<std::map<int, double>> ilu_S;                 // rows of the sparse matrix
  std::vector<std::vector<std::pair<int, double>>> ilu_LU;  // factored rows (L strictly below, U from the diagonal)
  // weights of one face: rule collapsed at P[0] (GW5x5 if singular or near, else GW1x1); singular (the
  // point is P[0]): single layer only. G_k = sum J w N_k / r, H_k = -sum (R . cr) w N_k / r^3
  static void ilu_face_rule(const V3 P[3], const V3& x, bool singular, bool near, double G[3], double H[3]) {
    static const double gw1[1][3] = {{0.5, 0.5, 1.0}};
    const V3 cr = cross(P[1] - P[0], P[2] - P[0]);
    const double J = norm(cr);
    for (int k = 0; k < 3; ++k) G[k] = H[k] = 0.;
    const bool fine = singular || near;
    const int nq = fine ? 25 : 1;
    for (int g = 0; g < nq; ++g) {
      const double t0 = fine ? kGW5[g][0] : gw1[g][0];
      const double t1 = fine ? kGW5[g][1] : gw1[g][1];
      const double ww = fine ? kGW5[g][2] : gw1[g][2];
      const double N[3] = {t0, t1 * (1. - t0), (1. - t0) * (1. - t1)};
      const V3 Xq = N[0] * P[0] + N[1] * P[1] + N[2] * P[2];
      const V3 R = Xq - x;
      const double r = norm(R), w = ww * (1. - t0);
      for (int k = 0; k < 3; ++k) {
        G[k] += J * w / r * N[k];
        if (!singular) H[k] -= dot(R, cr) * w / (r * r * r) * N[k];
      }
    }
  }
  void build_ilu(int neighborhood, int kring) {
    // faces per point
    std::vector<std::vector<int>> pfaces(n_nodes);
    for (int f = 0; f < n_faces; ++f)
      for (int a = 0; a < 3; ++a) pfaces[F[f].v[a]].push_back(f);
    // BUCKETS grid: cube of half-side 1.05 x the largest extent around the bounding-box centre,
    // clamp(round(cbrt(faces)), 8, 64) cells across; each face at its centroid, vertices, edge midpoints
    Bucket grid;
    std::vector<std::vector<int>> cell_faces;
    if (neighborhood == 0) {
      double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
      for (const V3& x : Xn)
        for (int a = 0; a < 3; ++a) {
          lo[a] = std::min(lo[a], x[a]);
          hi[a] = std::max(hi[a], x[a]);
        }
      const double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]});
      double ulo[3], uhi[3];
      for (int a = 0; a < 3; ++a) {
        ulo[a] = (lo[a] + hi[a]) / 2. - 1.05 * ext;
        uhi[a] = (lo[a] + hi[a]) / 2. + 1.05 * ext;
      }
      const int ncell = std::clamp((int)std::llround(std::cbrt((double)std::max(n_faces, 1))), 8, 64);
      grid.init(ulo, uhi, (2.1 * ext) / ncell);
      cell_faces.assign(grid.size[0] * grid.size[1] * grid.size[2], {});
      for (int f = 0; f < n_faces; ++f) {
        const V3* X = F[f].X;
        const V3 pts[7] = {F[f].centroid, X[0], X[1], X[2], 0.5 * (X[0] + X[1]), 0.5 * (X[1] + X[2]),
                           0.5 * (X[2] + X[0])};
        for (const V3& q : pts) {
          auto& cf = cell_faces[grid.cell_of(q)];
          if (std::find(cf.begin(), cf.end(), f) == cf.end()) cf.push_back(f);
        }
      }
    }
    // K-RING: faces across each edge
    std::map<std::pair<int, int>, std::vector<int>> edge_faces;
    if (neighborhood == 1)
      for (int f = 0; f < n_faces; ++f)
        for (int a = 0; a < 3; ++a) {
          int u = F[f].v[a], v = F[f].v[(a + 1) % 3];
          edge_faces[{std::min(u, v), std::max(u, v)}].push_back(f);
        }
    const double near_switch = scale / 10.;
    ilu_S.assign(n_dof, {});
    for (int k = 0; k < n_nodes; ++k) {
      if (point_dofs[k].empty()) continue;
      std::set<int> faces(pfaces[k].begin(), pfaces[k].end());
      if (neighborhood == 0) {
        for (int f : cell_faces[grid.cell_of(Xn[k])]) faces.insert(f);
      } else {
        std::set<int> front = faces;
        for (int ring = 0; ring < kring; ++ring) {
          std::set<int> next;
          for (int f : front)
            for (int a = 0; a < 3; ++a) {
              int u = F[f].v[a], v = F[f].v[(a + 1) % 3];
              for (int g : edge_faces[{std::min(u, v), std::max(u, v)}])
                if (!faces.count(g)) next.insert(g);
            }
          faces.insert(next.begin(), next.end());
          front = next;
        }
      }
      std::map<int, double> coeffs;
      for (int f : faces) {
        const FaceRec& fr = F[f];
        const int s = fr.v[0] == k ? 0 : fr.v[1] == k ? 1 : 2;  // getPoints(origin)
        const int ord[3] = {s, (s + 1) % 3, (s + 2) % 3};
        const V3 P[3] = {fr.X[ord[0]], fr.X[ord[1]], fr.X[ord[2]]};
        const bool singular = fr.v[ord[0]] == k;
        const bool near = norm(fr.centroid - Xn[k]) < near_switch;
        double G[3], H[3];
        ilu_face_rule(P, Xn[k], singular, near, G, H);
        for (int a = 0; a < 3; ++a) {
          const int col = fr.dof[ord[a]];
          const double val = dof_kind[col] == 0 ? G[a] : -H[a];
          if (val != 0.) coeffs[col] += val;
        }
      }
      for (int d : point_dofs[k]) {
        if (dof_kind[d] == 2) {
          ilu_S[d] = {{d, 1.0}};
          continue;
        }
        ilu_S[d] = coeffs;
        if (dof_kind[d] != 0) ilu_S[d][d] -= alpha[d];
        else if (!ilu_S[d].count(d)) ilu_S[d][d] = 1.0;
      }
    }
    ilu_LU.assign(n_dof, {});
    std::vector<std::map<int, double>> W = ilu_S;
    if (ilu_kind == 2) {
      // ILUT (ILUT_CRS, basic_linear_systems.hpp:2610-2760): fill-in allowed, then entries below
      // drop_tol * |row|_1 dropped and the largest max_entries of each of the L and U parts kept
      for (int i = 0; i < n_dof; ++i) {
        std::map<int, double> wi = ilu_S[i];
        double norm1 = 0.;
        for (const auto& [c, v] : ilu_S[i]) norm1 += std::abs(v);
        for (auto it = wi.begin(); it != wi.end() && it->first < i; ++it) {
          const int k = it->first;
          double dk = 0.;
          for (const auto& [c, v] : W[k])
            if (c == k) dk = v;
          if (!(std::abs(dk) >= ilut_pivot_min)) throw std::runtime_error("oracle ILUT: bad pivot");
          it->second /= dk;
          for (const auto& [c, v] : W[k])
            if (c > k) wi[c] -= it->second * v;
        }
        double dg = wi.count(i) ? wi[i] : 0.;
        if (std::abs(dg) < ilut_pivot_min) dg = dg >= 0. ? ilut_pivot_min : -ilut_pivot_min;
        std::vector<std::pair<int, double>> lo, up;
        for (const auto& [c, v] : wi) {
          if (c == i || v == 0. || (ilut_drop_tol > 0. && std::abs(v) < ilut_drop_tol * norm1)) continue;
          (c < i ? lo : up).push_back({c, v});
        }
        for (auto* part : {&lo, &up}) {
          if ((int)part->size() > ilut_max_entries) {
            std::sort(part->begin(), part->end(),
                      [](const auto& a, const auto& b) { return std::abs(a.second) > std::abs(b.second); });
            part->resize(ilut_max_entries);
          }
        }
        W[i].clear();
        for (const auto& e : lo) W[i][e.first] = e.second;
        W[i][i] = dg;
        for (const auto& e : up) W[i][e.first] = e.second;
        ilu_LU[i].assign(W[i].begin(), W[i].end());
      }
      return;
    }
    // ILU(0): row i in turn, each of its entries k < i (ascending) divided by U_kk and its multiple of row k's
    // upper part subtracted where row i has an entry (MILU(0): elsewhere subtracted from the diagonal, times omega)
    for (int i = 0; i < n_dof; ++i) {
      auto& wi = W[i];
      for (auto it = wi.begin(); it != wi.end() && it->first < i; ++it) {
        const int k = it->first;
        const double piv = W[k].at(k);
        if (!(std::abs(piv) >= 1e-20)) throw std::runtime_error("oracle ILU(0): bad pivot");
        it->second /= piv;
        for (auto jt = W[k].upper_bound(k); jt != W[k].end(); ++jt) {
          auto e = wi.find(jt->first);
          if (e != wi.end()) e->second -= it->second * jt->second;
          else if (ilu_kind == 1) wi.at(i) -= milu_omega * (it->second * jt->second);
        }
      }
      ilu_LU[i].assign(wi.begin(), wi.end());
    }
  }
  int ilu_kind = 0;  // 0 ILU(0), 1 MILU(0), 2 ILUT
  double milu_omega = 1.0, ilut_drop_tol = 1e-3, ilut_pivot_min = 1e-12;
  int ilut_max_entries = 50;
  void ilu_apply(const double* r, double* z) const {
    for (int i = 0; i < n_dof; ++i) {
      double s = r[i];
      for (const auto& [j, v] : ilu_LU[i])
        if (j < i) s -= v * z[j];
      z[i] = s;
    }
    for (int i = n_dof - 1; i >= 0; --i) {
      double s = z[i], d = 0.;
      for (const auto& [j, v] : ilu_LU[i]) {
        if (j > i) s -= v * z[j];
        if (j == i) d = v;
      }
      z[i] = s / d;
    }
  }

  // the reference's SCHWARZ on the same matrix (bem/BEM_solveBVP_GMRES_FMM.hpp:901-1013,
  // lib/include/basic_linear_systems.hpp:2800-3003): blocks from breadth-first searches over the matrix graph,
  // each block solved by LU with partial pivoting, the block solutions averaged over the blocks covering a DOF
  std::vector<std::vector<int>> swr_blocks;
  std::vector<int> swr_ncore;
  void build_schwarz_ref(int core_k, int overlap_k, int max_core, int max_block) {
    swr_blocks.clear();
    swr_ncore.clear();
    std::vector<bool> taken(n_dof, false);
    for (int seed = 0; seed < n_dof; ++seed) {
      if (taken[seed]) continue;
      std::vector<int> core;
      std::set<int> seen{seed};
      std::deque<std::pair<int, int>> q{{seed, 0}};
      while (!q.empty() && (int)core.size() < max_core) {
        auto [u, dep] = q.front();
        q.pop_front();
        if (!taken[u]) {
          taken[u] = true;
          core.push_back(u);
          if ((int)core.size() >= max_core) break;
        }
        if (dep >= core_k) continue;
        for (const auto& [v, val] : ilu_S[u]) {
          (void)val;
          if (v == u || seen.count(v)) continue;
          seen.insert(v);
          if (!taken[v]) q.push_back({v, dep + 1});
        }
      }
      if (core.empty()) {
        taken[seed] = true;
        core.push_back(seed);
      }
      std::vector<int> block = core;
      std::set<int> seen2(core.begin(), core.end());
      std::deque<std::pair<int, int>> q2;
      for (int u : core) q2.push_back({u, 0});
      while (!q2.empty() && (int)block.size() < max_block) {
        auto [u, dep] = q2.front();
        q2.pop_front();
        if (dep >= overlap_k) continue;
        for (const auto& [v, val] : ilu_S[u]) {
          (void)val;
          if (v == u || seen2.count(v)) continue;
          seen2.insert(v);
          block.push_back(v);
          if ((int)block.size() >= max_block) break;
          q2.push_back({v, dep + 1});
        }
      }
      swr_ncore.push_back((int)core.size());
      swr_blocks.push_back(block);
    }
  }
  double s_entry(int i, int j) const {
    auto it = ilu_S[i].find(j);
    return it == ilu_S[i].end() ? 0.0 : it->second;
  }
  void schwarz_ref_apply(const double* r, double* z) const {
    std::vector<double> sum(n_dof, 0.);
    std::vector<int> cover(n_dof, 0);
    for (const auto& blk : swr_blocks) {
      const int m = (int)blk.size();
      std::vector<double> A((size_t)m * m), y(m);
      for (int i = 0; i < m; ++i) {
        y[i] = r[blk[i]];
        for (int j = 0; j < m; ++j) A[(size_t)i * m + j] = s_entry(blk[i], blk[j]);
      }
      for (int k = 0; k < m; ++k) {  // LU with partial pivoting (lapack_lu's dgetrf/dgetrs)
        int pr = k;
        for (int i = k + 1; i < m; ++i)
          if (std::abs(A[(size_t)i * m + k]) > std::abs(A[(size_t)pr * m + k])) pr = i;
        if (pr != k) {
          for (int j = 0; j < m; ++j) std::swap(A[(size_t)k * m + j], A[(size_t)pr * m + j]);
          std::swap(y[k], y[pr]);
        }
        for (int i = k + 1; i < m; ++i) {
          const double f = A[(size_t)i * m + k] / A[(size_t)k * m + k];
          for (int j = k + 1; j < m; ++j) A[(size_t)i * m + j] -= f * A[(size_t)k * m + j];
          y[i] -= f * y[k];
        }
      }
      for (int i = m - 1; i >= 0; --i) {
        double acc = y[i];
        for (int j = i + 1; j < m; ++j) acc -= A[(size_t)i * m + j] * y[j];
        y[i] = acc / A[(size_t)i * m + i];
      }
      for (int i = 0; i < m; ++i) {
        sum[blk[i]] += y[i];
        cover[blk[i]] += 1;
      }
    }
    for (int g = 0; g < n_dof; ++g) z[g] = cover[g] ? sum[g] / cover[g] : 0.;
  }

  int precond = 0;  // 0 none, 1 Jacobi, 2 restricted additive Schwarz (blocks set beforehand), 3 ILU(0) (built),
                    // 4 the reference's SCHWARZ (built)
  int solve(const double* x0, double* xout, double tol, int max_iter, int restart, int* iters_out, double* res_out) {
    std::vector<double> b(n_dof);
    rhs(b.data());
    const std::vector<double> dinv = precond == 1 ? diag_inv() : std::vector<double>(n_dof, 1.0);
    std::vector<double> tmp(n_dof);
    auto apply_m = [&](std::vector<double>& v) {
      if (precond == 2) {
        schwarz_apply(v.data(), tmp.data());
        v.swap(tmp);
      } else if (precond == 3) {
        ilu_apply(v.data(), tmp.data());
        v.swap(tmp);
      } else if (precond == 4) {
        schwarz_ref_apply(v.data(), tmp.data());
        v.swap(tmp);
      } else {
        for (int i = 0; i < n_dof; ++i) v[i] *= dinv[i];
      }
    };
    apply_m(b);
    auto Av = [&](const std::vector<double>& v) {
      std::vector<double> y(n_dof);
      matvec(v.data(), y.data());
      apply_m(y);
      return y;
    };
    auto nrm = [](const std::vector<double>& v) {
      double s = 0;
      for (double a : v) s += a * a;
      return std::sqrt(s);
    };
    std::vector<double> x(x0, x0 + n_dof);
    for (int i = 0; i < n_dof; ++i)
      if (dof_kind[i] == 0) x[i] *= Lscale;  // phi_n unknowns are solved scaled, as xout is unscaled below
    int iters = 0;
    double cur_err = 1e300;
    const int restart_max = std::max(1, restart);
    while (iters < max_iter && cur_err > tol) {
      const int remaining = max_iter - iters;
      const int k_cycle_max = std::min(restart_max, remaining);
      // gmres(A, b, x, 1): r0 = b - A x; Arnoldi with one vector
      std::vector<double> r0 = Av(x);
      for (int i = 0; i < n_dof; ++i) r0[i] = b[i] - r0[i];
      const double beta = nrm(r0);
      std::vector<std::vector<double>> V;
      std::vector<std::vector<double>> H;  // H[i][j]
      std::vector<double> cs, sn, g;
      std::vector<std::vector<double>> Rm;
      const std::vector<double> xbase = x;
      auto add_basis = [&](const std::vector<double>& vlast) {
        std::vector<double> w = Av(vlast);
        const size_t nb = V.size();
        std::vector<double> hcol(nb + 1, 0.);
        for (size_t i = 0; i < nb; ++i) {
          double h = 0;
          for (int a = 0; a < n_dof; ++a) h += V[i][a] * w[a];
          hcol[i] = h;
          for (int a = 0; a < n_dof; ++a) w[a] -= h * V[i][a];
        }
        const double hn = nrm(w);
        hcol[nb] = hn;
        for (int a = 0; a < n_dof; ++a) w[a] /= hn;
        V.push_back(w);
        return hcol;
      };
      auto givens_col = [&](std::vector<double> col) {
        const size_t c = cs.size();
        for (size_t i = 0; i < c; ++i) {
          const double tmp = std::fma(cs[i], col[i], sn[i] * col[i + 1]);
          col[i + 1] = std::fma(cs[i], col[i + 1], -sn[i] * col[i]);
          col[i] = tmp;
        }
        double cc = 1.0, ss = 0.0;
        const double r = std::hypot(col[c], col[c + 1]);
        if (r != 0.0) {
          cc = col[c] / r;
          ss = col[c + 1] / r;
        }
        const double tmp = std::fma(cc, col[c], ss * col[c + 1]);
        col[c + 1] = std::fma(cc, col[c + 1], -ss * col[c]);
        col[c] = tmp;
        cs.push_back(cc);
        sn.push_back(ss);
        g.push_back(0.);
        const double gt = std::fma(cc, g[c], ss * g[c + 1]);
        g[c + 1] = std::fma(cc, g[c + 1], -ss * g[c]);
        g[c] = gt;
        Rm.push_back(col);
      };
      auto update_solution = [&]() {
        x = xbase;
        if (beta == 0.) return;
        const size_t n = cs.size();
        std::vector<double> y(n, 0.);
        for (int i = (int)n - 1; i >= 0; --i) {
          double s = g[i];
          for (size_t j = i + 1; j < n; ++j) s -= Rm[j][i] * y[j];
          y[i] = s / Rm[i][i];
        }
        for (size_t i = 0; i < n; ++i)
          for (int a = 0; a < n_dof; ++a) x[a] += y[i] * V[i][a];
      };
      std::vector<double> v0(n_dof);
      for (int a = 0; a < n_dof; ++a) v0[a] = r0[a] / beta;
      V.push_back(v0);
      g.assign(1, beta);
      givens_col(add_basis(v0));
      iters += 1;
      double est = std::abs(g.back());
      int sat = 0;
      while (true) {
        givens_col(add_basis(V.back()));
        ++iters;
        est = std::abs(g.back());
        if (est <= tol * 10.) {
          update_solution();
          std::vector<double> r = Av(x);
          for (int a = 0; a < n_dof; ++a) r[a] -= b[a];
          cur_err = nrm(r);
          if (cur_err <= tol) break;
          if (sat++ >= 5) break;
        }
        if (iters >= max_iter && (int)cs.size() >= k_cycle_max) {
          update_solution();
          if (est <= tol * 10.) {
            std::vector<double> r = Av(x);
            for (int a = 0; a < n_dof; ++a) r[a] -= b[a];
            cur_err = nrm(r);
          } else
            cur_err = est;
          break;
        }
      }
    }
    // unscale phin on Dirichlet DOFs (GMRES_FMM.hpp:3462-3486)
    for (int i = 0; i < n_dof; ++i) xout[i] = (dof_kind[i] == 0) ? x[i] / Lscale : x[i];
    if (iters_out) *iters_out = iters;
    if (res_out) *res_out = cur_err;
    return cur_err <= tol ? 0 : 1;
  }
};

}  // namespace

extern "C" {

struct oracle_stats {
  int n_levels;
  int n_buckets;
  int n_leaves;
  long long n_m2l_pairs;
  long long near_nnz;
  double L;
  double scale;
};

// TriShape<6> and its t0 / t1 derivatives at (t0, t1), 6 values each
void oracle_tri6_shape(double t0, double t1, double* N, double* D0, double* D1) { q6_shape(t0, t1, N, D0, D1); }

// npf = 6: true quadratic faces (tri and face_dof with six entries per face)
void* oracle_create_npf(int n_nodes, const double* X, int n_faces, const int* tri, const int* face_dof, int n_dof,
                        const int* dof_kind, const int* dof_point, const double* phi, const double* phin, int p,
                        double theta, int leaf_max, int max_level, int q, int npf) {
  try {
    if (npf != 3 && npf != 6) return nullptr;
    auto* o = new Oracle();
    o->npf = npf;
    o->n_nodes = n_nodes;
    o->n_faces = n_faces;
    o->n_dof = n_dof;
    o->P = p;
    o->theta = theta;
    o->leaf_max = leaf_max;
    o->max_level = max_level;
    o->q = q;
    o->setup(X, tri, face_dof, dof_kind, dof_point, phi, phin);
    return o;
  } catch (...) {
    return nullptr;
  }
}
void* oracle_create(int n_nodes, const double* X, int n_faces, const int* tri, const int* face_dof, int n_dof,
                    const int* dof_kind, const int* dof_point, const double* phi, const double* phin, int p,
                    double theta, int leaf_max, int max_level, int q) {
  return oracle_create_npf(n_nodes, X, n_faces, tri, face_dof, n_dof, dof_kind, dof_point, phi, phin, p, theta,
                           leaf_max, max_level, q, 3);
}
void oracle_destroy(void* h) { delete static_cast<Oracle*>(h); }
void oracle_set_robin(void* h, const double* kappa) { static_cast<Oracle*>(h)->set_robin(kappa); }
void oracle_matvec(void* h, const double* x, double* y) { static_cast<Oracle*>(h)->matvec(x, y); }
void oracle_rhs(void* h, double* b) { static_cast<Oracle*>(h)->rhs(b); }
void oracle_alpha(void* h, double* a) {
  auto* o = static_cast<Oracle*>(h);
  std::copy(o->alpha.begin(), o->alpha.end(), a);
}
void oracle_dense(void* h, double* A) { static_cast<Oracle*>(h)->dense(A); }
int oracle_solve(void* h, const double* x0, double* x, double tol, int max_iter, int restart, int* iters, double* res,
                 int precond) {
  static_cast<Oracle*>(h)->precond = precond;
  return static_cast<Oracle*>(h)->solve(x0, x, tol, max_iter, restart, iters, res);
}
void oracle_get_stats(void* h, oracle_stats* s) {
  auto* o = static_cast<Oracle*>(h);
  s->n_levels = (int)o->level_buckets.size();
  int nb = 0;
  for (auto& lv : o->level_buckets) nb += (int)lv.size();
  s->n_buckets = nb;
  s->n_leaves = (int)o->leaves.size();
  s->n_m2l_pairs = o->n_m2l_pairs;
  s->near_nnz = o->nnz;
  s->L = o->Lscale;
  s->scale = o->scale;
}
void oracle_set_m2l_cache(void* h, int on) {
  auto* o = static_cast<Oracle*>(h);
  o->m2l_cached = on != 0;
  o->m2l_ops.clear();
}
void oracle_set_schwarz(void* h, int n_blocks, const int64_t* ptr, const int32_t* idx, const int32_t* ncore) {
  auto* o = static_cast<Oracle*>(h);
  o->sw_ptr.assign(ptr, ptr + n_blocks + 1);
  o->sw_idx.assign(idx, idx + ptr[n_blocks]);
  o->sw_ncore.assign(ncore, ncore + n_blocks);
}
void oracle_schwarz_apply(void* h, const double* r, double* z) { static_cast<Oracle*>(h)->schwarz_apply(r, z); }
// the reference's sparse preconditioner matrix and its ILU(0); returns its entry count (-1 on failure)
long long oracle_ilu_build(void* h, int neighborhood, int kring) {
  auto* o = static_cast<Oracle*>(h);
  try {
    o->build_ilu(neighborhood, kring);
  } catch (...) {
    return -1;
  }
  long long nnz = 0;
  for (const auto& r : o->ilu_S) nnz += (long long)r.size();
  return nnz;
}
// unfactored rows: ptr (n + 1), col, val
void oracle_ilu_rows(void* h, long long* ptr, int* col, double* val) {
  auto* o = static_cast<Oracle*>(h);
  long long q = 0;
  ptr[0] = 0;
  for (int i = 0; i < o->n_dof; ++i) {
    for (const auto& [j, v] : o->ilu_S[i]) {
      col[q] = j;
      val[q] = v;
      ++q;
    }
    ptr[i + 1] = q;
  }
}
void oracle_ilu_apply(void* h, const double* r, double* z) { static_cast<Oracle*>(h)->ilu_apply(r, z); }
// the reference's SCHWARZ on the ILU(0) matrix (oracle_ilu_build first): returns the number of blocks
int oracle_schwarz_ref_build(void* h, int core_k, int overlap_k, int max_core, int max_block) {
  auto* o = static_cast<Oracle*>(h);
  o->build_schwarz_ref(core_k, overlap_k, max_core, max_block);
  return (int)o->swr_blocks.size();
}
void oracle_schwarz_ref_blocks(void* h, long long* ptr, int* idx, int* ncore) {
  auto* o = static_cast<Oracle*>(h);
  long long q = 0;
  ptr[0] = 0;
  for (size_t b = 0; b < o->swr_blocks.size(); ++b) {
    for (int g : o->swr_blocks[b]) idx[q++] = g;
    ptr[b + 1] = q;
    ncore[b] = o->swr_ncore[b];
  }
}
void oracle_schwarz_ref_apply(void* h, const double* r, double* z) { static_cast<Oracle*>(h)->schwarz_ref_apply(r, z); }
// factorisation used by oracle_ilu_build: 0 ILU(0), 1 MILU(0) (omega), 2 ILUT (drop_tol, max_entries, pivot_min)
void oracle_ilu_kind(void* h, int kind, double omega, double drop_tol, int max_entries, double pivot_min) {
  auto* o = static_cast<Oracle*>(h);
  o->ilu_kind = kind;
  o->milu_omega = omega;
  o->ilut_drop_tol = drop_tol;
  o->ilut_max_entries = max_entries;
  o->ilut_pivot_min = pivot_min;
}
long long oracle_ilu_factored_nnz(void* h) {
  long long q = 0;
  for (const auto& r : static_cast<Oracle*>(h)->ilu_LU) q += (long long)r.size();
  return q;
}
// factored rows: ptr (n + 1), col, val (ILU(0)/MILU(0): the pattern of the unfactored rows)
void oracle_ilu_factored(void* h, long long* ptr, int* col, double* val) {
  auto* o = static_cast<Oracle*>(h);
  long long q = 0;
  ptr[0] = 0;
  for (int i = 0; i < o->n_dof; ++i) {
    for (const auto& [j, v] : o->ilu_LU[i]) {
      col[q] = j;
      val[q] = v;
      ++q;
    }
    ptr[i + 1] = q;
  }
}
void oracle_ilu_face_rule(const double* P9, const double* x, int singular, int near, double* G, double* H) {
  const V3 P[3] = {{P9[0], P9[1], P9[2]}, {P9[3], P9[4], P9[5]}, {P9[6], P9[7], P9[8]}};
  Oracle::ilu_face_rule(P, {x[0], x[1], x[2]}, singular != 0, near != 0, G, H);
}
void oracle_diag_inv(void* h, double* d) {
  const auto v = static_cast<Oracle*>(h)->diag_inv();
  std::copy(v.begin(), v.end(), d);
}
// near-field row of a target node (sorted DOF indices) — for structural parity
int oracle_near_row(void* h, int t, int* idx_out, double* wg, double* wgn, int cap) {
  auto* o = static_cast<Oracle*>(h);
  const int n = (int)o->near_idx[t].size();
  for (int i = 0; i < std::min(n, cap); ++i) {
    idx_out[i] = o->near_idx[t][i];
    wg[i] = o->near_wG[t][i];
    wgn[i] = o->near_wGn[t][i];
  }
  return n;
}

// Harmonics probe for known-answer tests: Y_n^m (reference convention) at x, all |m|<=n<=P.
void oracle_sph_harmonics(const double* x, int P, double* out_re, double* out_im) {
  Harm h;
  h.eval({x[0], x[1], x[2]}, P);
  for (int i = 0; i < (P + 1) * (P + 1); ++i) {
    out_re[i] = h.Y[i].real();
    out_im[i] = h.Y[i].imag();
  }
}
// grad(R_n^q).nhat for all (n,q)
void oracle_dR(const double* x, const double* nh, int P, double* out_re, double* out_im) {
  SolidR S;
  S.eval({x[0], x[1], x[2]}, P);
  for (int n = 0; n <= P; ++n)
    for (int q = -n; q <= n; ++q) {
      const cd v = S.dR(n, q, {nh[0], nh[1], nh[2]});
      out_re[idx(n, q)] = v.real();
      out_im[idx(n, q)] = v.imag();
    }
}
// Point-source FMM invariant: potential of unit charges q_i at y_i (and dipoles d_i along nh_i),
// expanded P2M about c1, translated M2M to c2, M2L to c3, L2L to c4, evaluated L2P at x.
// Returns (single-layer, double-layer) potentials.
void oracle_translate_chain(int ns, const double* ys, const double* qs, const double* nhs, const double* c1,
                            const double* c2, const double* c3, const double* c4, const double* x, int P, double* out) {
  Tables T;
  T.build(P);
  const int C = (P + 1) * (P + 1);
  std::vector<cd> M0(C, 0.), M1(C, 0.);
  SolidR S;
  for (int s = 0; s < ns; ++s) {
    const V3 y = {ys[3 * s], ys[3 * s + 1], ys[3 * s + 2]};
    const V3 nh = {nhs[3 * s], nhs[3 * s + 1], nhs[3 * s + 2]};
    S.eval(y - V3{c1[0], c1[1], c1[2]}, P);
    for (int n = 0; n <= P; ++n)
      for (int m = -n; m <= n; ++m) {
        M0[idx(n, m)] += qs[s] * S.R[idx(n, -m)];
        M1[idx(n, m)] += qs[s] * S.dR(n, -m, nh);
      }
  }
  // M2M c1 -> c2
  std::vector<cd> N0(C, 0.), N1(C, 0.);
  S.eval(V3{c1[0] - c2[0], c1[1] - c2[1], c1[2] - c2[2]}, P);
  for (int j = 0; j <= P; ++j)
    for (int k = -j; k <= j; ++k)
      for (int n = 0; n <= j; ++n)
        for (int m = -n; m <= n; ++m) {
          if (std::abs(k - m) > j - n) continue;
          const cd a = T.m2m[(size_t)idx(j, k) * C + idx(n, m)] * S.R[idx(n, -m)];
          N0[idx(j, k)] += a * M0[idx(j - n, k - m)];
          N1[idx(j, k)] += a * M1[idx(j - n, k - m)];
        }
  // M2L c2 -> c3
  std::vector<cd> L0(C, 0.), L1(C, 0.);
  Harm H;
  H.eval(V3{c2[0] - c3[0], c2[1] - c3[1], c2[2] - c3[2]}, 2 * P);
  for (int j = 0; j <= P; ++j)
    for (int k = -j; k <= j; ++k)
      for (int n = 0; n <= P; ++n)
        for (int m = -n; m <= n; ++m) {
          const double rl = 1.0 / (std::pow(H.rho, j + n) * H.rho);
          const cd a = T.m2l[(size_t)idx(j, k) * C + idx(n, m)] * H.Y[idx(j + n, m - k)] * rl;
          L0[idx(j, k)] += a * N0[idx(n, m)];
          L1[idx(j, k)] += a * N1[idx(n, m)];
        }
  // L2L c3 -> c4
  std::vector<cd> K0(C, 0.), K1(C, 0.);
  S.eval(V3{c3[0] - c4[0], c3[1] - c4[1], c3[2] - c4[2]}, P);
  for (int j = 0; j <= P; ++j)
    for (int k = -j; k <= j; ++k)
      for (int n = j; n <= P; ++n)
        for (int m = -n; m <= n; ++m) {
          if (std::abs(m - k) > n - j) continue;
          const cd a = T.l2l[(size_t)idx(j, k) * C + idx(n, m)] * S.R[idx(n - j, m - k)];
          K0[idx(j, k)] += a * L0[idx(n, m)];
          K1[idx(j, k)] += a * L1[idx(n, m)];
        }
  S.eval(V3{x[0] - c4[0], x[1] - c4[1], x[2] - c4[2]}, P);
  cd f0 = 0., f1 = 0.;
  for (int n = 0; n <= P; ++n)
    for (int m = -n; m <= n; ++m) {
      f0 += S.R[idx(n, m)] * K0[idx(n, m)];
      f1 += S.R[idx(n, m)] * K1[idx(n, m)];
    }
  out[0] = f0.real();
  out[1] = f1.real();
  out[2] = f0.imag();
  out[3] = f1.imag();
}

}  // extern "C"
