// Near-field quadrature of one linear (or true quadratic) face for one target point, shared by
// the GPU kernel (kernels.hip) and host test code.
//
// fill_direct_linear / fill_nonadj_linear (reference bem/BEM_solveBVP_GMRES_FMM.hpp:1824-1891):
//   adjacent face (target is corner `slot`): 25-point GW5x5 rule collapsed at the target,
//     x(t0,t1) = t0 X_s + t1(1-t0) X_{s+1} + (1-t0)(1-t1) X_{s+2}, weight ww (1-t0) J
//     (ModTriShape<3>, basic_arithmetic_array_operations.hpp:1591-1600); the double-layer
//     weight of the face's corner 0 is dropped, as in the reference (acc.add(idx0, ., 0.0));
//   non-adjacent face with |centroid - x| < near_region: same GW5x5 rule collapsed at corner 0;
//   otherwise: the q-point Dunavant rule.
//   WG_k += J w N_k / r,  WGn_k += -(R/r . cr) w N_k / r^2,  R = x_q - x_target, cr = J n.
#pragma once

#ifdef __HIPCC__
#define BEMFMM_HD __host__ __device__
#else
#define BEMFMM_HD
#endif

namespace bemfmm {

// 1 / sqrt(a) for a > 0. On the GPU: the hardware estimate (v_rsq_f64, about 23 correct bits) refined by two
// Newton steps y += y (1 - a y^2) / 2, each of which doubles the correct bits, to within an ulp or two of the
// correctly rounded value; it replaces a correctly rounded square root followed by a division, the two longest
// instruction sequences of the quadrature loop. On the host: the correctly rounded expression.
BEMFMM_HD inline double inv_sqrt(double a) {
#if defined(__HIP_DEVICE_COMPILE__)
  double y = __builtin_amdgcn_rsq(a);
  y = fma(0.5 * y, fma(-(a * y), y, 1.0), y);
  y = fma(0.5 * y, fma(-(a * y), y, 1.0), y);
  return y;
#else
  return 1. / sqrt(a);
#endif
}

BEMFMM_HD inline void face_weights(const double* Xf, double J, const double* cr, const double* cent, int slot, double x,
                                   double y, double z, double near_region, const double (*gw5)[3],
                                   const double (*dun)[4], int q, double WG[3], double WGn[3]) {
  WG[0] = WG[1] = WG[2] = 0.;
  WGn[0] = WGn[1] = WGn[2] = 0.;
  const double crx = cr[0], cry = cr[1], crz = cr[2];
  if (slot >= 0) {
    const int p0 = slot, p1 = (slot + 1) % 3, p2 = (slot + 2) % 3;
    double G0 = 0., G1 = 0., G2 = 0., H0 = 0., H1 = 0., H2 = 0.;
    for (int g = 0; g < 25; ++g) {
      const double t0 = gw5[g][0], t1 = gw5[g][1], ww = gw5[g][2];
      const double N0 = t0, N1 = t1 * (1. - t0), N2 = (t0 - 1.) * (t1 - 1.);
      const double qx = N0 * Xf[3 * p0] + N1 * Xf[3 * p1] + N2 * Xf[3 * p2] - x;
      const double qy = N0 * Xf[3 * p0 + 1] + N1 * Xf[3 * p1 + 1] + N2 * Xf[3 * p2 + 1] - y;
      const double qz = N0 * Xf[3 * p0 + 2] + N1 * Xf[3 * p1 + 2] + N2 * Xf[3 * p2 + 2] - z;
      const double nr_inv = inv_sqrt(qx * qx + qy * qy + qz * qz);
      const double tmp = ww * (1. - t0) * nr_inv;
      const double rc = -((qx * nr_inv) * crx + (qy * nr_inv) * cry + (qz * nr_inv) * crz);
      G0 += J * (tmp * N0);
      G1 += J * (tmp * N1);
      G2 += J * (tmp * N2);
      H0 += rc * (tmp * N0) * nr_inv;
      H1 += rc * (tmp * N1) * nr_inv;
      H2 += rc * (tmp * N2) * nr_inv;
    }
    WG[p0] = G0;
    WG[p1] = G1;
    WG[p2] = G2;
    WGn[p0] = H0;
    WGn[p1] = H1;
    WGn[p2] = H2;
    WGn[0] = 0.;
    return;
  }
  const double dcx = cent[0] - x, dcy = cent[1] - y, dcz = cent[2] - z;
  if (sqrt(dcx * dcx + dcy * dcy + dcz * dcz) < near_region) {
    for (int g = 0; g < 25; ++g) {
      const double t0 = gw5[g][0], t1 = gw5[g][1], ww = gw5[g][2];
      const double N[3] = {t0, t1 * (1. - t0), (t0 - 1.) * (t1 - 1.)};
      const double qx = N[0] * Xf[0] + N[1] * Xf[3] + N[2] * Xf[6] - x;
      const double qy = N[0] * Xf[1] + N[1] * Xf[4] + N[2] * Xf[7] - y;
      const double qz = N[0] * Xf[2] + N[1] * Xf[5] + N[2] * Xf[8] - z;
      const double nr_inv = inv_sqrt(qx * qx + qy * qy + qz * qz);
      const double rc = (qx * nr_inv) * crx + (qy * nr_inv) * cry + (qz * nr_inv) * crz;
      for (int k = 0; k < 3; ++k) {
        const double tmp = (ww * (1. - t0) * N[k]) * nr_inv;
        WG[k] += J * tmp;
        WGn[k] -= rc * (tmp * nr_inv);
      }
    }
  } else {
    for (int g = 0; g < q; ++g) {
      const double N[3] = {dun[g][0], dun[g][1], dun[g][2]};
      const double ww = dun[g][3];
      const double qx = N[0] * Xf[0] + N[1] * Xf[3] + N[2] * Xf[6] - x;
      const double qy = N[0] * Xf[1] + N[1] * Xf[4] + N[2] * Xf[7] - y;
      const double qz = N[0] * Xf[2] + N[1] * Xf[5] + N[2] * Xf[8] - z;
      const double nr_inv = inv_sqrt(qx * qx + qy * qy + qz * qz);
      const double rc = (qx * nr_inv) * crx + (qy * nr_inv) * cry + (qz * nr_inv) * crz;
      for (int k = 0; k < 3; ++k) {
        const double tmp = ww * N[k] * nr_inv;
        WG[k] += J * tmp;
        WGn[k] -= rc * (tmp * nr_inv);
      }
    }
  }
}

// ---- true quadratic elements (6-node faces: vertices 0..2, edge midpoints 3 = (0,1), 4 = (1,2), 5 = (2,0)) ----
// TriShape<6> and its t0 / t1 derivatives (basic_arithmetic_array_operations.hpp:1467-1566), t2 = 1 - t0 - t1
BEMFMM_HD inline void tri6_shape(double t0, double t1, double N[6]) {
  const double t2 = 1. - t0 - t1;
  N[0] = t0 * (2. * t0 - 1.);
  N[1] = t1 * (2. * t1 - 1.);
  N[2] = t2 * (2. * t2 - 1.);
  N[3] = 4. * t0 * t1;
  N[4] = 4. * t1 * t2;
  N[5] = 4. * t2 * t0;
}
BEMFMM_HD inline void tri6_dshape(double t0, double t1, double D0[6], double D1[6]) {
  const double t2 = 1. - t0 - t1;
  D0[0] = -1. + 4. * t0; D0[1] = 0.; D0[2] = 1. - 4. * t2; D0[3] = 4. * t1; D0[4] = -4. * t1; D0[5] = -4. * t0 + 4. * t2;
  D1[0] = 0.; D1[1] = -1 + 4. * t1; D1[2] = 1. - 4. * t2; D1[3] = 4 * t0; D1[4] = 4 * t2 - 4 * t1; D1[5] = -4 * t0;
}
// point and tangent cross product (dX/dt0 x dX/dt1) of the 6-node geometry X6 (18 values) at (t0, t1)
BEMFMM_HD inline void tri6_geom(const double* X6, double t0, double t1, double x[3], double cr[3]) {
  double N[6], D0[6], D1[6];
  tri6_shape(t0, t1, N);
  tri6_dshape(t0, t1, D0, D1);
  double a[3] = {0., 0., 0.}, b[3] = {0., 0., 0.};
  for (int c = 0; c < 3; ++c) x[c] = 0.;
  for (int k = 0; k < 6; ++k)
    for (int c = 0; c < 3; ++c) {
      x[c] += N[k] * X6[3 * k + c];
      a[c] += D0[k] * X6[3 * k + c];
      b[c] += D1[k] * X6[3 * k + c];
    }
  cr[0] = a[1] * b[2] - a[2] * b[1];
  cr[1] = a[2] * b[0] - a[0] * b[2];
  cr[2] = a[0] * b[1] - a[1] * b[0];
}
// trueQuadN6 (Network.hpp:1607-1632): the weight of an edge midpoint without a DOF (bit e of `active` clear)
// goes half to each end vertex
BEMFMM_HD inline void tri6_density(double t0, double t1, unsigned active, double N[6]) {
  tri6_shape(t0, t1, N);
  for (int e = 0; e < 3; ++e)
    if (!((active >> e) & 1u)) {
      N[e] += 0.5 * N[3 + e];
      N[(e + 1) % 3] += 0.5 * N[3 + e];
      N[3 + e] = 0.;
    }
}

// Near-field weights of one true quadratic face for one target (fill_direct_true_quad / fill_nonadj_true_quad,
// BEM_solveBVP_GMRES_FMM.hpp:1424-1592): `slot` 0..2 = the target is that vertex (GW5x5 collapsed at it, its own
// double-layer weight dropped), 3..5 = the target is that edge midpoint (GW5x5 on the two halves of the face split
// at the midpoint, its own double-layer weight dropped), -1 = not adjacent: the Dunavant rule `dun_near` (q
// points) if the vertex centroid lies within near_region of the target, else the 6-point rule `dun6`.
// WG_k += |cr| w N_k / r, WGn_k += -(R/r . cr) w N_k / r^2, R = x_q - x_target, N the redistributed TriShape<6>.
BEMFMM_HD inline void face_weights_quad(const double* X6, unsigned active, const double* cent, int slot, double x,
                                        double y, double z, double near_region, const double (*gw5)[3],
                                        const double (*dun_near)[4], int q, const double (*dun6)[4], double WG[6],
                                        double WGn[6]) {
  for (int k = 0; k < 6; ++k) WG[k] = WGn[k] = 0.;
  auto point = [&](double b0, double b1, double w, int drop) {
    double xq[3], cr[3], N[6];
    tri6_geom(X6, b0, b1, xq, cr);
    tri6_density(b0, b1, active, N);
    const double Rx = xq[0] - x, Ry = xq[1] - y, Rz = xq[2] - z;
    const double ir = inv_sqrt(Rx * Rx + Ry * Ry + Rz * Rz);
    const double J = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    const double g = J * w * ir;
    const double gn = -((Rx * ir) * cr[0] + (Ry * ir) * cr[1] + (Rz * ir) * cr[2]) * w * ir * ir;
    for (int k = 0; k < 6; ++k) {
      WG[k] += g * N[k];
      if (k != drop) WGn[k] += gn * N[k];
    }
  };
  if (slot >= 0 && slot < 3) {
    const int v0 = slot, v1 = (slot + 1) % 3, v2 = (slot + 2) % 3;
    for (int g = 0; g < 25; ++g) {
      const double t0 = gw5[g][0], t1 = gw5[g][1], ww = gw5[g][2];
      double bary[3];
      bary[v0] = t0;
      bary[v1] = t1 * (1. - t0);
      bary[v2] = (1. - t0) * (1. - t1);
      point(bary[0], bary[1], ww * (1. - t0), slot);
    }
    return;
  }
  if (slot >= 3) {
    const int e = slot - 3, va = e, vb = (e + 1) % 3, vc = (e + 2) % 3;
    for (int g = 0; g < 25; ++g) {
      const double t0 = gw5[g][0], t1 = gw5[g][1], ww = gw5[g][2];
      const double omt0 = 1. - t0, s1 = t1 * omt0, s2 = omt0 * (1. - t1);
      double bary[3];
      bary[va] = 0.5 * t0;
      bary[vb] = 0.5 * t0 + s1;
      bary[vc] = s2;
      point(bary[0], bary[1], ww * omt0 * 0.5, slot);
      bary[va] = 0.5 * t0 + s2;
      bary[vb] = 0.5 * t0;
      bary[vc] = s1;
      point(bary[0], bary[1], ww * omt0 * 0.5, slot);
    }
    return;
  }
  const double dcx = cent[0] - x, dcy = cent[1] - y, dcz = cent[2] - z;
  const bool near = sqrt(dcx * dcx + dcy * dcy + dcz * dcz) < near_region;
  const double(*rule)[4] = near ? dun_near : dun6;
  const int n = near ? q : 6;
  for (int g = 0; g < n; ++g) point(rule[g][0], rule[g][1], rule[g][3], -1);
}

}  // namespace bemfmm
