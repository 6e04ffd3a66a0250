// Element-level device math shared by the stiffness, plastic-body and
// batched kernels: GLL tables, hex8 Jacobians, B-matrix derivatives,
// isotropic elasticity, stress invariants and the Mohr-Coulomb viscoplastic
// point update.  Every function follows the reference routine it cites, in
// the reference's operation order; the including file disables contraction.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sgeo {

// l_m'(xi_q) of the 1D Lagrange basis at the GLL nodes {-1,0,1}
// (src/gll_library.f90:257-299; exact dyadic values)
__device__ __forceinline__ double gll_d1(int q, int m) {
  constexpr double d1[3][3] = {{-1.5, 2.0, -0.5}, {-0.5, 0.0, 0.5},
                               {0.5, -2.0, 1.5}};
  return d1[q][m];
}

// gll_weights(i) = wx(ix)*wy(iy)*wz(iz), wx = {1/3, 4/3, 1/3} as zwgljd
// returns them for np = 3 (src/gll_library.f90:52)
__device__ __forceinline__ double gll_weight(int i) {
  constexpr double wx[3] = {1.0 / 3.0, 4.0 / 3.0, 1.0 / 3.0};
  return wx[i % 3] * wx[(i / 3) % 3] * wx[i / 9];
}

// dshape_hex8(a, g, i) at GLL point i (src/shape_library.f90:157-182);
// xi/eta/zeta of point i are -1, 0, 1 so every value is exact.
__device__ __forceinline__ void dshape_point(int i, double ds[3][8]) {
  const double x[3] = {-1.0, 0.0, 1.0};
  const double e8 = 0.125;
  const double xp = 1.0 + x[i % 3], xm = 1.0 - x[i % 3];
  const double ep = 1.0 + x[(i / 3) % 3], em = 1.0 - x[(i / 3) % 3];
  const double zp = 1.0 + x[i / 9], zm = 1.0 - x[i / 9];
  ds[0][0] = -e8 * em * zm; ds[0][1] = e8 * em * zm;
  ds[0][2] = e8 * ep * zm;  ds[0][3] = -e8 * ep * zm;
  ds[0][4] = -e8 * em * zp; ds[0][5] = e8 * em * zp;
  ds[0][6] = e8 * ep * zp;  ds[0][7] = -e8 * ep * zp;
  ds[1][0] = -e8 * xm * zm; ds[1][1] = -e8 * xp * zm;
  ds[1][2] = e8 * xp * zm;  ds[1][3] = e8 * xm * zm;
  ds[1][4] = -e8 * xm * zp; ds[1][5] = -e8 * xp * zp;
  ds[1][6] = e8 * xp * zp;  ds[1][7] = e8 * xm * zp;
  ds[2][0] = -e8 * xm * em; ds[2][1] = -e8 * xp * em;
  ds[2][2] = -e8 * xp * ep; ds[2][3] = -e8 * xm * ep;
  ds[2][4] = e8 * xm * em;  ds[2][5] = e8 * xp * em;
  ds[2][6] = e8 * xp * ep;  ds[2][7] = e8 * xm * ep;
}

// Jacobian at GLL point i, its determinant and inverse
// (src/slope.F90:324-326; src/math_library.f90:78-80,108-125).
// xc[g*3+d]: corner g, axis d.  jinv[a][b].
__device__ __forceinline__ double jacobian_point(int i, const double* xc,
                                                 double jinv[3][3]) {
  double ds[3][8];
  dshape_point(i, ds);
  double j[3][3];
  for (int b = 0; b < 3; b++)
    for (int a = 0; a < 3; a++) {
      double s = 0.0;
      for (int g = 0; g < 8; g++) s = s + ds[a][g] * xc[g * 3 + b];
      j[a][b] = s;
    }
  double det = j[0][0] * (j[1][1] * j[2][2] - j[2][1] * j[1][2]);
  det = det - j[0][1] * (j[1][0] * j[2][2] - j[2][0] * j[1][2]);
  det = det + j[0][2] * (j[1][0] * j[2][1] - j[2][0] * j[1][1]);
  const double j11 = j[1][1] * j[2][2] - j[2][1] * j[1][2];
  const double j21 = -j[1][0] * j[2][2] + j[2][0] * j[1][2];
  const double j31 = j[1][0] * j[2][1] - j[2][0] * j[1][1];
  const double j12 = -j[0][1] * j[2][2] + j[2][1] * j[0][2];
  const double j22 = j[0][0] * j[2][2] - j[2][0] * j[0][2];
  const double j32 = -j[0][0] * j[2][1] + j[2][0] * j[0][1];
  const double j13 = j[0][1] * j[1][2] - j[1][1] * j[0][2];
  const double j23 = -j[0][0] * j[1][2] + j[1][0] * j[0][2];
  const double j33 = j[0][0] * j[1][1] - j[1][0] * j[0][1];
  jinv[0][0] = j11 / det; jinv[0][1] = j12 / det; jinv[0][2] = j13 / det;
  jinv[1][0] = j21 / det; jinv[1][1] = j22 / det; jinv[1][2] = j23 / det;
  jinv[2][0] = j31 / det; jinv[2][1] = j32 / det; jinv[2][2] = j33 / det;
  return det;
}

// deriv(a, n) = sum_b jinv(a,b) * dlagrange_gll(b, i, n) (src/slope.F90:328).
// dlagrange_gll(b,i,n) is l'(..) of the b axis when n and i agree in the two
// other indices, else 0 (exact products with 1 and 0,
// src/gll_library.f90:58-80).
__device__ __forceinline__ void deriv_node(int i, int n, const double jinv[3][3],
                                           double d[3]) {
  const int ix = i % 3, iy = (i / 3) % 3, iz = i / 9;
  const int nx = n % 3, ny = (n / 3) % 3, nz = n / 9;
  const double dl0 = (iy == ny && iz == nz) ? gll_d1(ix, nx) : 0.0;
  const double dl1 = (ix == nx && iz == nz) ? gll_d1(iy, ny) : 0.0;
  const double dl2 = (ix == nx && iy == ny) ? gll_d1(iz, nz) : 0.0;
  for (int a = 0; a < 3; a++) {
    double s = 0.0;
    s = s + jinv[a][0] * dl0;
    s = s + jinv[a][1] * dl1;
    s = s + jinv[a][2] * dl2;
    d[a] = s;
  }
}

// compute_cmat (src/elastic.f90:30-39): the distinct entries
struct Cmat {
  double c11, c12, c44;
};
__device__ __forceinline__ Cmat cmat_of(double E, double v) {
  const double onepv = 1.0 + v, onemv = 1.0 - v, onem2v = 1.0 - 2.0 * v;
  const double Efact = E / (onepv * onem2v);
  Cmat c;
  c.c11 = onemv * Efact;
  c.c12 = v * Efact;
  c.c44 = 0.5 * onem2v * Efact;
  return c;
}

// sigma = matmul(cmat, eps) with the zero entries of cmat skipped (adding an
// exact zero product never changes a sum)
__device__ __forceinline__ void cmat_apply(const Cmat& c, const double* e,
                                           double* s) {
  s[0] = c.c11 * e[0] + c.c12 * e[1] + c.c12 * e[2];
  s[1] = c.c12 * e[0] + c.c11 * e[1] + c.c12 * e[2];
  s[2] = c.c12 * e[0] + c.c12 * e[1] + c.c11 * e[2];
  s[3] = c.c44 * e[3];
  s[4] = c.c44 * e[4];
  s[5] = c.c44 * e[5];
}

// stress_invariant (src/math_library.f90:160-184)
__device__ __forceinline__ void stress_invariant(const double* s, double& sigm,
                                                 double& dsbar, double& theta) {
  const double small = 1.e-12;
  const double sq3 = sqrt(3.0);
  const double s1 = s[0], s2 = s[1], s3 = s[2], s4 = s[3], s5 = s[4], s6 = s[5];
  sigm = (s1 + s2 + s3) / 3.0;
  double d2 = ((s1 - s2) * (s1 - s2) + (s2 - s3) * (s2 - s3) +
               (s3 - s1) * (s3 - s1)) / 6.0 + s4 * s4 + s5 * s5 + s6 * s6;
  if (d2 < small) d2 = small;
  const double ds1 = s1 - sigm, ds2 = s2 - sigm, ds3 = s3 - sigm;
  const double d3 = ds1 * ds2 * ds3 - ds1 * s5 * s5 - ds2 * s6 * s6 -
                    ds3 * s4 * s4 + 2.0 * s4 * s5 * s6;
  dsbar = sq3 * sqrt(d2);
  if (dsbar < small) {
    theta = 0.0;
  } else {
    const double sd = sqrt(d2);
    double sine = -3.0 * sq3 * d3 / (2.0 * (sd * sd * sd));
    if (sine >= 1.0) sine = 1.0;
    if (sine < -1.0) sine = -1.0;
    theta = asin(sine) / 3.0;
  }
}

// mohcouf (src/plastic_library.f90:101-106): the Mohr-Coulomb yield value;
// phi in degrees, theta in radians
__device__ __forceinline__ double mohcouf(double phi, double c, double sigm, double dsbar,
                                          double theta) {
  const double deg2rad = 3.141592653589793 / 180.0;
  const double phir = phi * deg2rad;
  double snph, csph, snth, csth;
  sincos(phir, &snph, &csph);
  sincos(theta, &snth, &csth);
  return snph * sigm + dsbar * (csth / sqrt(3.0) - snth * snph / 3.0) - c * csph;
}

// mohcouq (src/plastic_library.f90:123-141): derivatives of the potential
// with respect to the three invariants; psi in degrees
__device__ __forceinline__ void mohcouq(double psi, double dsbar, double theta,
                                        double& dq1, double& dq2, double& dq3) {
  const double deg2rad = 3.141592653589793 / 180.0;
  const double psir = psi * deg2rad;
  const double snth = sin(theta);
  const double snps = sin(psir);
  const double sq3 = sqrt(3.0);
  dq1 = snps;
  if (fabs(snth) > 0.49) {
    const double c1 = snth < 0.0 ? -1.0 : 1.0;
    dq2 = (sq3 * 0.5 - c1 * snps * 0.5 / sq3) * sq3 * 0.5 / dsbar;
    dq3 = 0.0;
  } else {
    const double csth = cos(theta);
    const double cs3th = cos(3.0 * theta);
    const double tn3th = tan(3.0 * theta);
    const double tnth = snth / csth;
    dq2 = sq3 * csth / dsbar * ((1.0 + tnth * tn3th) + snps * (tn3th - tnth) / sq3) * 0.5;
    dq3 = 0.5 * 3.0 * (sq3 * snth + snps * csth) / (cs3th * dsbar * dsbar);
  }
}

// Factored material block of one strength-reduction trial
// (src/plastic_library.f90:42-88) plus the trial's pseudo-time step.
struct MatParams {
  double ym, nuf, phif, cohf, psif, dt;
};

// Strain, stress, yield and viscoplastic update at one GLL point
// (src/slope.F90:329-377; src/excavation.F90:561-611).  ue: element
// displacement (81).  Updates ev (evpt, 6) in place; returns f and writes
// eff (effective stress), devp = C evp (zero when f < 0) and, when inv is
// given, the invariants sigm, dsbar, theta of eff.
__device__ __forceinline__ double point_update(int i, const double jinv[3][3],
                                               const double* ue,
                                               const MatParams& mm,
                                               const double* st, double* ev,
                                               double* eff, double* devp,
                                               double* inv = nullptr) {
  const Cmat cm = cmat_of(mm.ym, mm.nuf);
  double eps[6] = {0, 0, 0, 0, 0, 0};
  for (int n = 0; n < 27; n++) {  // eps = matmul(bmat, eld)
    double d[3];
    deriv_node(i, n, jinv, d);
    const double ux = ue[3 * n], uy = ue[3 * n + 1], uz = ue[3 * n + 2];
    eps[0] = eps[0] + d[0] * ux;
    eps[1] = eps[1] + d[1] * uy;
    eps[2] = eps[2] + d[2] * uz;
    eps[3] = eps[3] + d[1] * ux;
    eps[3] = eps[3] + d[0] * uy;
    eps[4] = eps[4] + d[2] * uy;
    eps[4] = eps[4] + d[1] * uz;
    eps[5] = eps[5] + d[2] * ux;
    eps[5] = eps[5] + d[0] * uz;
  }
  double sig[6];
  for (int s = 0; s < 6; s++) eps[s] = eps[s] - ev[s];
  cmat_apply(cm, eps, sig);
  for (int s = 0; s < 6; s++) eff[s] = sig[s] + st[s];
  double sigm, dsbar, theta;
  stress_invariant(eff, sigm, dsbar, theta);
  if (inv) {  // the invariants of the effective stress, for the caller
    inv[0] = sigm;
    inv[1] = dsbar;
    inv[2] = theta;
  }
  const double f = mohcouf(mm.phif, mm.cohf, sigm, dsbar, theta);
  for (int s = 0; s < 6; s++) devp[s] = 0.0;
  if (!(f >= 0.0)) return f;
  double dq1, dq2, dq3;
  mohcouq(mm.psif, dsbar, theta, dq1, dq2, dq3);
  // formm (src/plastic_library.f90:164-206), flow = f*(m1 dq1+m2 dq2+m3 dq3),
  // erate = matmul(flow, effsigma), evp = erate*dt (src/slope.F90:351-356)
  const double sx = eff[0], sy = eff[1], sz = eff[2];
  const double txy = eff[3], tyz = eff[4], tzx = eff[5];
  const double sm = (sx + sy + sz) / 3.0;
  const double dx = sx - sm, dy = sy - sm, dz = sz - sm;
  const double m1v = 1.0 / (3.0 * sm);
  double evp[6];
  for (int p = 0; p < 6; p++) {
    double acc = 0.0;
    for (int q = 0; q < 6; q++) {
      const int a = p < q ? p : q, b = p < q ? q : p;  // symmetric entries
      const double M1 = (a < 3 && b < 3) ? m1v : 0.0;
      double M2 = 0.0;
      if (a == b) M2 = a < 3 ? 2.0 : 6.0;
      else if (b < 3) M2 = -1.0;
      double M3;
      switch (a * 6 + b) {
        case 0: M3 = dx; break;          case 1: M3 = dz; break;
        case 2: M3 = dy; break;          case 3: M3 = txy; break;
        case 4: M3 = -2.0 * tyz; break;  case 5: M3 = tzx; break;
        case 7: M3 = dy; break;          case 8: M3 = dx; break;
        case 9: M3 = txy; break;         case 10: M3 = tyz; break;
        case 11: M3 = -2.0 * tzx; break; case 14: M3 = dz; break;
        case 15: M3 = -2.0 * txy; break; case 16: M3 = tyz; break;
        case 17: M3 = tzx; break;        case 21: M3 = -3.0 * dz; break;
        case 22: M3 = 3.0 * tzx; break;  case 23: M3 = 3.0 * tyz; break;
        case 28: M3 = -3.0 * dx; break;  case 29: M3 = 3.0 * txy; break;
        case 35: M3 = -3.0 * dy; break;  default: M3 = 0.0;
      }
      const double fl = f * ((M1 / 3.0) * dq1 + (M2 / 3.0) * dq2 +
                             (M3 / 3.0) * dq3);
      acc = acc + fl * eff[q];
    }
    evp[p] = acc * mm.dt;
  }
  for (int p = 0; p < 6; p++) ev[p] = ev[p] + evp[p];
  cmat_apply(cm, evp, devp);
  return f;
}

// eload(j) = matmul(devp, bmat)(j) for dof j = 3n + comp of node n
// (src/slope.F90:361); d = deriv(:, n) of the point
__device__ __forceinline__ double eload_dof(int comp, const double* devp,
                                            const double d[3]) {
  if (comp == 0) return devp[0] * d[0] + devp[3] * d[1] + devp[5] * d[2];
  if (comp == 1) return devp[1] * d[1] + devp[3] * d[0] + devp[4] * d[2];
  return devp[2] * d[2] + devp[4] * d[1] + devp[5] * d[0];
}

}  // namespace sgeo
