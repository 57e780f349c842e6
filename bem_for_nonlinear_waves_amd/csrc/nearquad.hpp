// Near-field quadrature of one linear triangle for one target point, shared by
// the GPU kernel (bemfmm.hip) and host test code.
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
      const double nr_inv = 1. / sqrt(qx * qx + qy * qy + qz * qz);
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
      const double nr_inv = 1. / sqrt(qx * qx + qy * qy + qz * qz);
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
      const double nr_inv = 1. / sqrt(qx * qx + qy * qy + qz * qz);
      const double rc = (qx * nr_inv) * crx + (qy * nr_inv) * cry + (qz * nr_inv) * crz;
      for (int k = 0; k < 3; ++k) {
        const double tmp = ww * N[k] * nr_inv;
        WG[k] += J * tmp;
        WGn[k] -= rc * (tmp * nr_inv);
      }
    }
  }
}

}  // namespace bemfmm
