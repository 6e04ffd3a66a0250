// sgeo_kernels.hip -- CDNA4 (gfx950) kernels of the EBE-PCG solve and the
// viscoplastic Mohr-Coulomb plastic-iteration body of SPECFEM3D_GEOTECH.
//
// Floating-point contraction is disabled in this file: every product and sum
// is rounded as written, in the order the reference evaluates it
// (src/solver.f90, src/slope.F90:312-384, src/preprocess.f90:108-144), so
// the element-level arithmetic matches the reference bit for bit; only the
// global dot products (tree reductions) and device transcendentals differ.
#pragma clang fp contract(off)

#include "sgeo_device.h"
#include "sgeo_internal.h"

namespace sgeo {

// --------------------------------------------------------------------------
// constant tables (uploaded once per device)
// --------------------------------------------------------------------------
// 1D derivative of the Lagrange basis at the GLL nodes {-1,0,1}:
// d1[q][m] = l_m'(xi_q)  (src/gll_library.f90:257-299; exact dyadic values)
__constant__ double c_d1[3][3];
// gll_weights(27) = wx*wy*wz (src/gll_library.f90:52)
__constant__ double c_w[27];

__host__ void upload_tables(hipStream_t s, const double* d1, const double* w) {
  hipMemcpyToSymbolAsync(HIP_SYMBOL(c_d1), d1, sizeof(double) * 9, 0,
                         hipMemcpyHostToDevice, s);
  hipMemcpyToSymbolAsync(HIP_SYMBOL(c_w), w, sizeof(double) * 27, 0,
                         hipMemcpyHostToDevice, s);
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

// deriv(a, n) = sum_b jinv(a,b) * dlagrange_gll(b, i, n)
// (src/slope.F90:328).  dlagrange_gll(b,i,n) is d1[i_b][n_b] when n and i
// agree in the two other indices, else 0 (exact products with 1 and 0).
__device__ __forceinline__ void deriv_node(int i, int n, const double jinv[3][3],
                                           double d[3]) {
  const int ix = i % 3, iy = (i / 3) % 3, iz = i / 9;
  const int nx = n % 3, ny = (n / 3) % 3, nz = n / 9;
  const double dl0 = (iy == ny && iz == nz) ? c_d1[ix][nx] : 0.0;
  const double dl1 = (ix == nx && iz == nz) ? c_d1[iy][ny] : 0.0;
  const double dl2 = (ix == nx && iy == ny) ? c_d1[iz][nz] : 0.0;
  for (int a = 0; a < 3; a++) {
    double s = 0.0;
    s = s + jinv[a][0] * dl0;
    s = s + jinv[a][1] * dl1;
    s = s + jinv[a][2] * dl2;
    d[a] = s;
  }
}

// compute_cmat (src/elastic.f90:30-39): the 4 distinct values
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

// --------------------------------------------------------------------------
// element stiffness: src/preprocess.f90:108-131
// --------------------------------------------------------------------------
// One block of 256 threads per element.  Thread t owns entries
// idx = t + 256*q of K_e (column-major, r = idx % 81, c = idx / 81).
// Per GLL point: T = B^T C (81x6) in LDS, then K += (T B)*detjac*w.
// Writes K (FP64, reference layout), the per-element diagonal and eld.
__global__ __launch_bounds__(256) void k_stiffness(
    int32_t nelmt, const double* __restrict__ corners,
    const int32_t* __restrict__ mat_id, const double* __restrict__ ym,
    const double* __restrict__ nu, double* __restrict__ K,
    double* __restrict__ kdiag, double* __restrict__ eld) {
  const int64_t e = blockIdx.x;
  if (e >= nelmt) return;  // uniform per block
  __shared__ double xc[24];
  __shared__ double deriv[3][27];
  __shared__ double T[81][6];
  __shared__ double s_det;
  const int t = threadIdx.x;
  if (t < 24) xc[t] = corners[e * 24 + t];
  const int m = mat_id[e] - 1;
  const Cmat cm = cmat_of(ym[m], nu[m]);
  // C(t,s) as a dense 6x6 for the T product
  double acc[26];
#pragma unroll
  for (int q = 0; q < 26; q++) acc[q] = 0.0;
  __syncthreads();
  for (int i = 0; i < 27; i++) {
    if (t < 27) {
      double jinv[3][3];
      const double det = jacobian_point(i, xc, jinv);
      double d[3];
      deriv_node(i, t, jinv, d);
      deriv[0][t] = d[0];
      deriv[1][t] = d[1];
      deriv[2][t] = d[2];
      if (t == 0) s_det = det;
    }
    __syncthreads();
    // T(r, s) = sum_t B(t,r) C(t,s)   (matmul(transpose(bmat), cmat))
    for (int q = t; q < 81 * 6; q += 256) {
      const int r = q % 81, s = q / 81;
      const int n = r / 3, comp = r % 3;
      const double dx = deriv[0][n], dy = deriv[1][n], dz = deriv[2][n];
      // nonzero B(t, r) for this column r (src/preprocess.f90:56-64)
      double b[6] = {0, 0, 0, 0, 0, 0};
      if (comp == 0) { b[0] = dx; b[3] = dy; b[5] = dz; }
      else if (comp == 1) { b[1] = dy; b[3] = dx; b[4] = dz; }
      else { b[2] = dz; b[4] = dy; b[5] = dx; }
      double C[6];
      if (s < 3) {
        C[0] = (s == 0) ? cm.c11 : cm.c12;
        C[1] = (s == 1) ? cm.c11 : cm.c12;
        C[2] = (s == 2) ? cm.c11 : cm.c12;
        C[3] = 0.0; C[4] = 0.0; C[5] = 0.0;
      } else {
        C[0] = 0.0; C[1] = 0.0; C[2] = 0.0;
        C[3] = (s == 3) ? cm.c44 : 0.0;
        C[4] = (s == 4) ? cm.c44 : 0.0;
        C[5] = (s == 5) ? cm.c44 : 0.0;
      }
      double v = 0.0;
#pragma unroll
      for (int tt = 0; tt < 6; tt++) v = v + b[tt] * C[tt];
      T[r][s] = v;
    }
    __syncthreads();
    const double det = s_det;
    const double w = c_w[i];
#pragma unroll
    for (int q = 0; q < 26; q++) {
      const int idx = t + 256 * q;
      if (idx < 6561) {
        const int r = idx % 81, c = idx / 81;
        const int n = c / 3, comp = c % 3;
        const double dx = deriv[0][n], dy = deriv[1][n], dz = deriv[2][n];
        // M(r,c) = sum_s T(r,s) B(s,c), ascending s over the nonzero B(s,c)
        double v;
        if (comp == 0) v = T[r][0] * dx + T[r][3] * dy + T[r][5] * dz;
        else if (comp == 1) v = T[r][1] * dy + T[r][3] * dx + T[r][4] * dz;
        else v = T[r][2] * dz + T[r][4] * dy + T[r][5] * dx;
        acc[q] = acc[q] + v * det * w;
      }
    }
    if (t == 0) eld[e * 81 + 3 * i + 2] = det * w;
    if (t == 1) { eld[e * 81 + 3 * i] = 0.0; eld[e * 81 + 3 * i + 1] = 0.0; }
    __syncthreads();
  }
  double* Ke = K + e * 6561;
#pragma unroll
  for (int q = 0; q < 26; q++) {
    const int idx = t + 256 * q;
    if (idx < 6561) {
      Ke[idx] = acc[q];
      const int r = idx % 81, c = idx / 81;
      if (r == c) kdiag[e * 81 + r] = acc[q];
    }
  }
}

// FP32 copy of the FP64 element matrices
__global__ void k_round_f32(int64_t n, const double* __restrict__ a,
                            float* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

// --------------------------------------------------------------------------
// element x GLL plastic body: src/slope.F90:312-383 (and the excavation
// variant src/excavation.F90:546-616 when excav != 0)
// --------------------------------------------------------------------------
#define BODY_EPB 9                 // elements per block
#define BODY_THREADS 256           // 9*27 = 243 point threads

struct PointOut {
  double jinv[9];
  double det;
  double devp[6];
};

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

__global__ __launch_bounds__(BODY_THREADS) void k_plastic_body(
    int32_t nel, const int32_t* __restrict__ elmt,  // NULL: identity
    const double* __restrict__ corners, const int32_t* __restrict__ gdof,
    const int32_t* __restrict__ mat_id, const MatParams* __restrict__ mp,
    const double* __restrict__ x, double* __restrict__ evpt,
    double* __restrict__ stress, double* __restrict__ BL,
    uint8_t* __restrict__ flags, PlasticScalars* __restrict__ ps,
    int is_last, int excav, const int32_t* __restrict__ g_num,
    double* __restrict__ vm_pt) {
  __shared__ double s_x[BODY_EPB][81];
  __shared__ double s_xc[BODY_EPB][24];
  __shared__ PointOut s_pt[BODY_EPB][27];
  __shared__ uint8_t s_flag[BODY_EPB][27];
  __shared__ double red[BODY_THREADS / 64];
  const int conv_or_last = ps->nl_isconv || is_last;
  const int t = threadIdx.x;
  const int64_t e0 = (int64_t)blockIdx.x * BODY_EPB;
  // stage element vectors and corners
  for (int q = t; q < BODY_EPB * 81; q += BODY_THREADS) {
    const int el = q / 81, j = q % 81;
    const int64_t a = e0 + el;
    if (a < nel) s_x[el][j] = x[gdof[a * 81 + j]];
  }
  for (int q = t; q < BODY_EPB * 24; q += BODY_THREADS) {
    const int el = q / 24, j = q % 24;
    const int64_t a = e0 + el;
    if (a < nel) {
      const int64_t e = elmt ? elmt[a] : a;
      s_xc[el][j] = corners[e * 24 + j];
    }
  }
  __syncthreads();
  double fmax = 0.0;
  int nplast = 0;
  const int el = t / 27, i = t % 27;
  const int64_t a = e0 + el;
  if (t < BODY_EPB * 27 && a < nel) {
    const int64_t e = elmt ? elmt[a] : a;
    const int m = mat_id[e] - 1;
    const MatParams mm = mp[m];
    const Cmat cm = cmat_of(mm.ym, mm.nuf);
    double jinv[3][3];
    const double det = jacobian_point(i, s_xc[el], jinv);
    // eps = matmul(bmat, eld) (src/slope.F90:329-330)
    double eps[6] = {0, 0, 0, 0, 0, 0};
    for (int n = 0; n < 27; n++) {
      double d[3];
      deriv_node(i, n, jinv, d);
      const double ux = s_x[el][3 * n], uy = s_x[el][3 * n + 1],
                   uz = s_x[el][3 * n + 2];
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
    double* ev = evpt + (e * 27 + i) * 6;
    double* st = stress + (e * 27 + i) * 6;
    double evl[6], sig[6], eff[6];
    for (int s = 0; s < 6; s++) evl[s] = ev[s];
    for (int s = 0; s < 6; s++) eps[s] = eps[s] - evl[s];
    cmat_apply(cm, eps, sig);
    for (int s = 0; s < 6; s++) eff[s] = sig[s] + st[s];
    double sigm, dsbar, theta;
    stress_invariant(eff, sigm, dsbar, theta);
    // mohcouf (src/plastic_library.f90:101-106)
    const double deg2rad = 3.141592653589793 / 180.0;
    const double phir = mm.phif * deg2rad;
    double snph, csph, snth, csth;
    sincos(phir, &snph, &csph);
    sincos(theta, &snth, &csth);
    const double f = snph * sigm + dsbar * (csth / sqrt(3.0) - snth * snph / 3.0) -
                     mm.cohf * csph;
    fmax = f > 0.0 ? f : 0.0;
    const bool plastic = f >= 0.0;
    if (flags) flags[e * 27 + i] = plastic ? 1 : 0;
    nplast = plastic ? 1 : 0;
    double devp[6] = {0, 0, 0, 0, 0, 0};
    if (plastic) {
      // mohcouq (src/plastic_library.f90:123-141)
      const double psir = mm.psif * deg2rad;
      const double snth2 = sin(theta);
      const double snps = sin(psir);
      const double sq3 = sqrt(3.0);
      const double dq1 = snps;
      double dq2, dq3;
      if (fabs(snth2) > 0.49) {
        const double c1 = snth2 < 0.0 ? -1.0 : 1.0;
        dq2 = (sq3 * 0.5 - c1 * snps * 0.5 / sq3) * sq3 * 0.5 / dsbar;
        dq3 = 0.0;
      } else {
        const double csth2 = cos(theta);
        const double cs3th = cos(3.0 * theta);
        const double tn3th = tan(3.0 * theta);
        const double tnth = snth2 / csth2;
        dq2 = sq3 * csth2 / dsbar *
              ((1.0 + tnth * tn3th) + snps * (tn3th - tnth) / sq3) * 0.5;
        dq3 = 0.5 * 3.0 * (sq3 * snth2 + snps * csth2) / (cs3th * dsbar * dsbar);
      }
      // formm (src/plastic_library.f90:164-206) and flow = f*(m1 dq1 + ...)
      const double sx = eff[0], sy = eff[1], sz = eff[2];
      const double txy = eff[3], tyz = eff[4], tzx = eff[5];
      const double sm = (sx + sy + sz) / 3.0;
      const double dx = sx - sm, dy = sy - sm, dz = sz - sm;
      double M1[6][6], M2[6][6], M3[6][6];
      for (int p = 0; p < 6; p++)
        for (int q = 0; q < 6; q++) { M1[p][q] = 0.0; M2[p][q] = 0.0; M3[p][q] = 0.0; }
      const double m1v = 1.0 / (3.0 * sm);
      for (int p = 0; p < 3; p++)
        for (int q = 0; q < 3; q++) M1[p][q] = m1v;
      for (int p = 0; p < 3; p++) { M2[p][p] = 2.0; M2[p + 3][p + 3] = 6.0; }
      M2[0][1] = -1.0; M2[0][2] = -1.0; M2[1][2] = -1.0;
      M3[0][0] = dx; M3[0][1] = dz; M3[0][2] = dy; M3[0][3] = txy;
      M3[0][4] = -2.0 * tyz; M3[0][5] = tzx;
      M3[1][1] = dy; M3[1][2] = dx; M3[1][3] = txy; M3[1][4] = tyz;
      M3[1][5] = -2.0 * tzx;
      M3[2][2] = dz; M3[2][3] = -2.0 * txy; M3[2][4] = tyz; M3[2][5] = tzx;
      M3[3][3] = -3.0 * dz; M3[3][4] = 3.0 * tzx; M3[3][5] = 3.0 * tyz;
      M3[4][4] = -3.0 * dx; M3[4][5] = 3.0 * txy;
      M3[5][5] = -3.0 * dy;
      for (int p = 0; p < 6; p++)
        for (int q = p + 1; q < 6; q++) {
          M1[q][p] = M1[p][q]; M2[q][p] = M2[p][q]; M3[q][p] = M3[p][q];
        }
      double evp[6];
      for (int p = 0; p < 6; p++) {
        double acc = 0.0;
        for (int q = 0; q < 6; q++) {
          const double fl = f * ((M1[p][q] / 3.0) * dq1 + (M2[p][q] / 3.0) * dq2 +
                                 (M3[p][q] / 3.0) * dq3);
          acc = acc + fl * eff[q];
        }
        evp[p] = acc * mm.dt;
      }
      for (int p = 0; p < 6; p++) { evl[p] = evl[p] + evp[p]; ev[p] = evl[p]; }
      cmat_apply(cm, evp, devp);
    }
    const bool contrib = plastic && !conv_or_last;
    s_flag[el][i] = contrib ? 1 : 0;
    PointOut& po = s_pt[el][i];
    po.det = det;
    for (int p = 0; p < 3; p++)
      for (int q = 0; q < 3; q++) po.jinv[p * 3 + q] = jinv[p][q];
    for (int p = 0; p < 6; p++) po.devp[p] = devp[p];
    if (conv_or_last) {
      if (excav && vm_pt) {
        double d = 0.0;
        for (int s = 0; s < 6; s++) d = d + evl[s] * evl[s];
        vm_pt[a * 27 + i] = sqrt((2.0 / 3.0) * d);
      }
      for (int s = 0; s < 6; s++) st[s] = eff[s];
    }
  }
  // block reductions of fmax / plastic count (order independent)
  const double bmax = block_max<BODY_THREADS / 64>(fmax, red);
  const double bcnt = block_sum<BODY_THREADS / 64>((double)nplast, red);
  if (t == 0) {
    atomicMax((unsigned long long*)&ps->fmax_bits,
              (unsigned long long)__double_as_longlong(bmax));
    atomicAdd(&ps->n_plastic, (int)bcnt);
  }
  if (conv_or_last) return;  // no body load (src/slope.F90:380)
  __syncthreads();
  // phase 2: bload(j) = sum_i (eload_i(j)*detjac_i)*w_i over the plastic
  // points, ascending i (src/slope.F90:361-362)
  for (int q = t; q < BODY_EPB * 81; q += BODY_THREADS) {
    const int el2 = q / 81, j = q % 81;
    const int64_t a2 = e0 + el2;
    if (a2 >= nel) continue;
    const int n = j / 3, comp = j % 3;
    double bl = 0.0;
    for (int i2 = 0; i2 < 27; i2++) {
      if (!s_flag[el2][i2]) continue;
      const PointOut& po = s_pt[el2][i2];
      double jinv[3][3];
      for (int p = 0; p < 3; p++)
        for (int r = 0; r < 3; r++) jinv[p][r] = po.jinv[p * 3 + r];
      double d[3];
      deriv_node(i2, n, jinv, d);
      double ed;
      if (comp == 0) ed = po.devp[0] * d[0] + po.devp[3] * d[1] + po.devp[5] * d[2];
      else if (comp == 1) ed = po.devp[1] * d[1] + po.devp[3] * d[0] + po.devp[4] * d[2];
      else ed = po.devp[2] * d[2] + po.devp[4] * d[1] + po.devp[5] * d[0];
      bl = bl + ed * po.det * c_w[i2];
    }
    BL[a2 * 81 + j] = bl;
  }
}

// dinv = 1/d (index 0 and zero entries -> 0), src/slope.F90:199
__global__ void k_invert_diag(int64_t n1, double* __restrict__ d) {
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n1;
       g += (int64_t)gridDim.x * blockDim.x)
    d[g] = (g == 0) ? 0.0 : 1.0 / d[g];
}

// Y[e][j] = eld[e][j] * gam[mat_id[e]-1]  (eld*gam of src/preprocess.f90:136)
__global__ void k_scale_by_mat(int32_t nelmt, const double* __restrict__ eld,
                               const int32_t* __restrict__ mat_id,
                               const double* __restrict__ gam,
                               double* __restrict__ Y) {
  const int64_t n = (int64_t)nelmt * 81;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x)
    Y[q] = eld[q] * gam[mat_id[q / 81] - 1];
}

// --------------------------------------------------------------------------
// launch wrappers
// --------------------------------------------------------------------------
static inline unsigned nblk(int64_t n, int per) {
  return (unsigned)((n + per - 1) / per);
}

unsigned vec_reduce_blocks(int64_t n1) {
  return nblk(n1, SGEO_VEC_THREADS * SGEO_VEC_ITEMS);
}

static inline unsigned grid_stride_blocks(int64_t n) {
  const int64_t b = (n + SGEO_VEC_THREADS - 1) / SGEO_VEC_THREADS;
  return (unsigned)(b < 65536 ? (b > 0 ? b : 1) : 65536);
}

void launch_stiffness(hipStream_t s, int32_t nelmt, const double* corners,
                      const int32_t* mat_id, const double* ym, const double* nu,
                      double* K, double* kdiag, double* eld) {
  if (nelmt <= 0) return;
  hipLaunchKernelGGL(k_stiffness, dim3(nelmt), dim3(256), 0, s, nelmt, corners,
                     mat_id, ym, nu, K, kdiag, eld);
}

void launch_round_f32(hipStream_t s, int64_t n, const double* a, float* b) {
  hipLaunchKernelGGL(k_round_f32, dim3(grid_stride_blocks(n)),
                     dim3(SGEO_VEC_THREADS), 0, s, n, a, b);
}

void launch_invert_diag(hipStream_t s, int64_t n1, double* d) {
  hipLaunchKernelGGL(k_invert_diag, dim3(grid_stride_blocks(n1)),
                     dim3(SGEO_VEC_THREADS), 0, s, n1, d);
}

void launch_scale_by_mat(hipStream_t s, int32_t nelmt, const double* eld,
                         const int32_t* mat_id, const double* gam, double* Y) {
  hipLaunchKernelGGL(k_scale_by_mat, dim3(grid_stride_blocks((int64_t)nelmt * 81)),
                     dim3(SGEO_VEC_THREADS), 0, s, nelmt, eld, mat_id, gam, Y);
}

void launch_plastic_body(hipStream_t s, int32_t nel, const int32_t* elmt,
                         const double* corners, const int32_t* gdof,
                         const int32_t* mat_id, const MatParams* mp,
                         const double* x, double* evpt, double* stress,
                         double* BL, uint8_t* flags, PlasticScalars* ps,
                         int is_last, int excav, const int32_t* g_num,
                         double* vm_pt) {
  if (nel <= 0) return;
  hipLaunchKernelGGL(k_plastic_body, dim3(nblk(nel, BODY_EPB)),
                     dim3(BODY_THREADS), 0, s, nel, elmt, corners, gdof,
                     mat_id, mp, x, evpt, stress, BL, flags, ps, is_last,
                     excav, g_num, vm_pt);
}

}  // namespace sgeo
