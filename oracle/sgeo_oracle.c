/*
 * sgeo_oracle.c -- CPU restatement of SPECFEM3D_GEOTECH's EBE-PCG solve and
 * viscoplastic Mohr-Coulomb plastic-iteration body (NGLL = 3).
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never the product).  Each function
 * follows the reference routine cited above it, in the reference's order of
 * floating-point operations wherever that order is fixed by the source
 * (matmul / dot_product are restated as sequential sums over the inner
 * index, element loops run in element order).  Built with
 * -ffp-contract=off so no fused multiply-adds appear that the reference's
 * x86-64 build does not have, and with -fno-builtin so that sin/cos pairs
 * are merged into sincos exactly where written (mohcouf), as the reference
 * build does.  OpenMP is used only where the result does not
 * depend on it: per-element products are computed in parallel into
 * per-element buffers, and every global accumulation then runs serially in
 * element order.
 */
#define _GNU_SOURCE /* sincos */
#include "sgeo_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NG 27
#define NED 81
#define NST 6

static const double PI_REF = 3.141592653589793;         /* src/global.f90:19 */
#define DEG2RAD (PI_REF / 180.0)                        /* src/global.f90:42 */
#define RAD2DEG (180.0 / PI_REF)

/* ------------------------------------------------------------------------ */
/* GLL tables                                                               */
/* ------------------------------------------------------------------------ */

/* lagrange1dGLL (src/gll_library.f90:257-299) at the 3 GLL nodes */
static void lagrange1d(const double *xii, double xi, double *phi,
                       double *dphi) {
  for (int i = 0; i < 3; i++) {
    double term[2], dterm[2];
    int k = 0;
    phi[i] = 1.0;
    for (int j = 0; j < 3; j++) {
      if (j == i) continue;
      term[k] = (xi - xii[j]) / (xii[i] - xii[j]);
      dterm[k] = 1.0 / (xii[i] - xii[j]);
      phi[i] = phi[i] * (xi - xii[j]) / (xii[i] - xii[j]);
      k++;
    }
    double sum_term[2] = {1.0, 1.0};
    for (int j = 0; j < 2; j++)
      for (int kk = 0; kk < 2; kk++)
        sum_term[j] = sum_term[j] * (kk == j ? dterm[kk] : term[kk]);
    dphi[i] = 0.0;
    for (int j = 0; j < 2; j++) dphi[i] = dphi[i] + sum_term[j];
  }
}

void orc_gll_tables(double *w, double *dlag, double *dsh) {
  /* zwgljd for np=3, alpha=beta=0: nodes -1, 0, 1; weights 1/3, 4/3, 1/3
   * (the values the reference computes, pinned by oracle/gll_probe.f90) */
  const double x[3] = {-1.0, 0.0, 1.0};
  const double wx[3] = {1.0 / 3.0, 4.0 / 3.0, 1.0 / 3.0};
  int n = 0;
  for (int k = 0; k < 3; k++)
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++) {
        w[n] = wx[i] * wx[j] * wx[k]; /* src/gll_library.f90:52 */
        n++;
      }
  /* dlagrange_gll(d, ii, n) (src/gll_library.f90:58-80) */
  for (int ii = 0; ii < NG; ii++) {
    double lx[3], ly[3], lz[3], dx[3], dy[3], dz[3];
    lagrange1d(x, x[ii % 3], lx, dx);
    lagrange1d(x, x[(ii / 3) % 3], ly, dy);
    lagrange1d(x, x[ii / 9], lz, dz);
    int m = 0;
    for (int k = 0; k < 3; k++)
      for (int j = 0; j < 3; j++)
        for (int i = 0; i < 3; i++) {
          dlag[0 + 3 * (ii + NG * m)] = dx[i] * ly[j] * lz[k];
          dlag[1 + 3 * (ii + NG * m)] = lx[i] * dy[j] * lz[k];
          dlag[2 + 3 * (ii + NG * m)] = lx[i] * ly[j] * dz[k];
          m++;
        }
  }
  /* dshape_function_hex8 (src/shape_library.f90:157-182) */
  const double e8 = 0.125;
  int ig = 0;
  for (int k = 0; k < 3; k++) {
    double zp = 1.0 + x[k], zm = 1.0 - x[k];
    for (int j = 0; j < 3; j++) {
      double ep = 1.0 + x[j], em = 1.0 - x[j];
      for (int i = 0; i < 3; i++) {
        double xp = 1.0 + x[i], xm = 1.0 - x[i];
        double *d = dsh + 24 * ig;
#define DS(a, g) d[(a) + 3 * (g)]
        DS(0, 0) = -e8 * em * zm; DS(0, 1) = e8 * em * zm;
        DS(0, 2) = e8 * ep * zm;  DS(0, 3) = -e8 * ep * zm;
        DS(0, 4) = -e8 * em * zp; DS(0, 5) = e8 * em * zp;
        DS(0, 6) = e8 * ep * zp;  DS(0, 7) = -e8 * ep * zp;
        DS(1, 0) = -e8 * xm * zm; DS(1, 1) = -e8 * xp * zm;
        DS(1, 2) = e8 * xp * zm;  DS(1, 3) = e8 * xm * zm;
        DS(1, 4) = -e8 * xm * zp; DS(1, 5) = -e8 * xp * zp;
        DS(1, 6) = e8 * xp * zp;  DS(1, 7) = e8 * xm * zp;
        DS(2, 0) = -e8 * xm * em; DS(2, 1) = -e8 * xp * em;
        DS(2, 2) = -e8 * xp * ep; DS(2, 3) = -e8 * xm * ep;
        DS(2, 4) = e8 * xm * em;  DS(2, 5) = e8 * xp * em;
        DS(2, 6) = e8 * xp * ep;  DS(2, 7) = e8 * xm * ep;
#undef DS
        ig++;
      }
    }
  }
}

typedef struct {
  double w[NG];
  double dlag[3 * NG * NG];
  double dsh[3 * 8 * NG];
} tables_t;

static const tables_t *tables(void) {
  static tables_t t;
  static int init = 0;
  if (!init) {
    orc_gll_tables(t.w, t.dlag, t.dsh);
    init = 1;
  }
  return &t;
}

/* ------------------------------------------------------------------------ */
/* small dense kernels                                                      */
/* ------------------------------------------------------------------------ */

void orc_compute_cmat(double E, double v, double *c) {
  /* src/elastic.f90:30-39, c(i,j) at c[i + 6*j] */
  memset(c, 0, sizeof(double) * 36);
  double onepv = 1.0 + v, onemv = 1.0 - v, onem2v = 1.0 - 2.0 * v;
  double Efact = E / (onepv * onem2v);
  c[0] = onemv * Efact;
  c[0 + 6 * 1] = v * Efact;
  c[0 + 6 * 2] = c[0 + 6 * 1];
  c[1 + 6 * 0] = c[0 + 6 * 1];
  c[1 + 6 * 1] = c[0];
  c[1 + 6 * 2] = c[0 + 6 * 1];
  c[2 + 6 * 0] = c[0 + 6 * 2];
  c[2 + 6 * 1] = c[1 + 6 * 2];
  c[2 + 6 * 2] = c[0];
  c[3 + 6 * 3] = 0.5 * onem2v * Efact;
  c[4 + 6 * 4] = c[3 + 6 * 3];
  c[5 + 6 * 5] = c[3 + 6 * 3];
}

/* j(a,b) at j[a + 3*b] */
#define J(a, b) j[(a) + 3 * (b)]
double orc_determinant3(const double *j) {
  /* src/math_library.f90:78-80 */
  double det = J(0, 0) * (J(1, 1) * J(2, 2) - J(2, 1) * J(1, 2));
  det = det - J(0, 1) * (J(1, 0) * J(2, 2) - J(2, 0) * J(1, 2));
  det = det + J(0, 2) * (J(1, 0) * J(2, 1) - J(2, 0) * J(1, 1));
  return det;
}

void orc_invert3(double *j) {
  /* src/math_library.f90:108-125 */
  double det = orc_determinant3(j);
  double j11 = J(1, 1) * J(2, 2) - J(2, 1) * J(1, 2);
  double j21 = -J(1, 0) * J(2, 2) + J(2, 0) * J(1, 2);
  double j31 = J(1, 0) * J(2, 1) - J(2, 0) * J(1, 1);
  double j12 = -J(0, 1) * J(2, 2) + J(2, 1) * J(0, 2);
  double j22 = J(0, 0) * J(2, 2) - J(2, 0) * J(0, 2);
  double j32 = -J(0, 0) * J(2, 1) + J(2, 0) * J(0, 1);
  double j13 = J(0, 1) * J(1, 2) - J(1, 1) * J(0, 2);
  double j23 = -J(0, 0) * J(1, 2) + J(1, 0) * J(0, 2);
  double j33 = J(0, 0) * J(1, 1) - J(1, 0) * J(0, 1);
  J(0, 0) = j11 / det; J(0, 1) = j12 / det; J(0, 2) = j13 / det;
  J(1, 0) = j21 / det; J(1, 1) = j22 / det; J(1, 2) = j23 / det;
  J(2, 0) = j31 / det; J(2, 1) = j32 / det; J(2, 2) = j33 / det;
}
#undef J

/* Jacobian, detJ, inverse and B-matrix data at GLL point i of an element:
 * jac = matmul(dshape_hex8(:,:,i), coord)  (coord(g,d) = corner g, axis d)
 * deriv = matmul(jac^-1, dlagrange_gll(:,i,:))
 * (src/slope.F90:324-329, src/preprocess.f90:116-121) */
static double gll_point_deriv(const tables_t *T, const double *xc, int i,
                              double *deriv /* deriv[a + 3*n] */) {
  double jac[9];
  const double *ds = T->dsh + 24 * i;
  for (int b = 0; b < 3; b++)
    for (int a = 0; a < 3; a++) {
      double s = 0.0;
      for (int g = 0; g < 8; g++) s = s + ds[a + 3 * g] * xc[b + 3 * g];
      jac[a + 3 * b] = s;
    }
  double detjac = orc_determinant3(jac);
  orc_invert3(jac);
  for (int n = 0; n < NG; n++) {
    const double *dl = T->dlag + 3 * (i + NG * n);
    for (int a = 0; a < 3; a++) {
      double s = 0.0;
      for (int b = 0; b < 3; b++) s = s + jac[a + 3 * b] * dl[b];
      deriv[a + 3 * n] = s;
    }
  }
  return detjac;
}

/* compute_bmat (src/preprocess.f90:47-65); bmat(s,j) at b[s + 6*j] */
static void compute_bmat(const double *deriv, double *b) {
  memset(b, 0, sizeof(double) * NST * NED);
  for (int n = 0; n < NG; n++) {
    int j1 = 3 * n, j2 = j1 + 1, j3 = j2 + 1;
    double dx = deriv[0 + 3 * n], dy = deriv[1 + 3 * n], dz = deriv[2 + 3 * n];
    b[0 + 6 * j1] = dx;
    b[1 + 6 * j2] = dy;
    b[2 + 6 * j3] = dz;
    b[3 + 6 * j1] = dy;
    b[3 + 6 * j2] = dx;
    b[4 + 6 * j2] = dz;
    b[4 + 6 * j3] = dy;
    b[5 + 6 * j1] = dz;
    b[5 + 6 * j3] = dx;
  }
}

/* ------------------------------------------------------------------------ */
/* stiffness (src/preprocess.f90:74-148)                                    */
/* ------------------------------------------------------------------------ */

static void element_km(const tables_t *T, const double *xc, double E, double nu,
                       double *km, double *eld) {
  double cmat[36], deriv[3 * NG], b[NST * NED], bt_c[NED * NST];
  orc_compute_cmat(E, nu, cmat);
  memset(km, 0, sizeof(double) * NED * NED);
  memset(eld, 0, sizeof(double) * NED);
  for (int i = 0; i < NG; i++) {
    double detjac = gll_point_deriv(T, xc, i, deriv);
    compute_bmat(deriv, b);
    /* matmul(transpose(bmat), cmat): bt_c(r,s) = sum_t b(t,r) c(t,s) */
    for (int s = 0; s < NST; s++)
      for (int r = 0; r < NED; r++) {
        double acc = 0.0;
        for (int t = 0; t < NST; t++) acc = acc + b[t + 6 * r] * cmat[t + 6 * s];
        bt_c[r + NED * s] = acc;
      }
    /* km = km + matmul(bt_c, bmat)*detjac*w */
    for (int c = 0; c < NED; c++)
      for (int r = 0; r < NED; r++) {
        double acc = 0.0;
        for (int s = 0; s < NST; s++) acc = acc + bt_c[r + NED * s] * b[s + 6 * c];
        km[r + NED * c] = km[r + NED * c] + acc * detjac * T->w[i];
      }
    /* eld(3*i) += detjac*w (src/preprocess.f90:123-125) */
    eld[3 * i + 2] = eld[3 * i + 2] + detjac * T->w[i];
  }
}

void orc_stiffness_bodyload(int32_t nelmt, int32_t neq, const double *xc,
                            const int32_t *gdof, const int32_t *mat_id,
                            const double *gam, const double *nu,
                            const double *ym, double *storekm, double *dprecon,
                            double *extload, const double *eqk) {
  const tables_t *T = tables();
  double *eld_all = (double *)malloc(sizeof(double) * NED * (size_t)nelmt);
#pragma omp parallel for schedule(static)
  for (int32_t e = 0; e < nelmt; e++) {
    int m = mat_id[e] - 1;
    element_km(T, xc + 24 * (size_t)e, ym[m], nu[m],
               storekm + (size_t)NED * NED * e, eld_all + (size_t)NED * e);
  }
  memset(dprecon, 0, sizeof(double) * ((size_t)neq + 1));
  for (int32_t e = 0; e < nelmt; e++) {
    const double *km = storekm + (size_t)NED * NED * e;
    const int32_t *g = gdof + (size_t)NED * e;
    for (int k = 0; k < NED; k++) dprecon[g[k]] = dprecon[g[k]] + km[k + NED * k];
    if (!extload) continue;
    const double *eld = eld_all + (size_t)NED * e;
    double gm = gam[mat_id[e] - 1];
    double tmp[NED];
    for (int k = 0; k < NED; k++) tmp[k] = extload[g[k]] - eld[k] * gm;
    for (int k = 0; k < NED; k++) extload[g[k]] = tmp[k];
    if (eqk) { /* src/preprocess.f90:137-143 */
      double eq[NED];
      for (int k = 0; k < NG; k++) {
        eq[3 * k + 0] = eqk[0] * eld[3 * k + 2];
        eq[3 * k + 1] = eqk[1] * eld[3 * k + 2];
        eq[3 * k + 2] = eqk[2] * eld[3 * k + 2];
      }
      for (int k = 0; k < NED; k++) tmp[k] = extload[g[k]] + eq[k] * gm;
      for (int k = 0; k < NED; k++) extload[g[k]] = tmp[k];
    }
  }
  dprecon[0] = 0.0;
  if (extload) extload[0] = 0.0;
  free(eld_all);
}

/* ------------------------------------------------------------------------ */
/* EBE-PCG (src/solver.f90:14-79)                                           */
/* ------------------------------------------------------------------------ */

void orc_ebe_matvec(int32_t neq, int32_t nelmt, const double *k,
                    const int32_t *gdof, const double *p, double *kp) {
  double *y = (double *)malloc(sizeof(double) * NED * (size_t)nelmt);
#pragma omp parallel for schedule(static)
  for (int32_t e = 0; e < nelmt; e++) {
    const double *km = k + (size_t)NED * NED * e;
    const int32_t *g = gdof + (size_t)NED * e;
    double pe[NED];
    for (int j = 0; j < NED; j++) pe[j] = p[g[j]];
    double *ye = y + (size_t)NED * e;
    /* matmul(km, p(egdof)): y(r) = sum_c km(r,c)*pe(c), c ascending */
    for (int r = 0; r < NED; r++) ye[r] = 0.0;
    for (int c = 0; c < NED; c++)
      for (int r = 0; r < NED; r++) ye[r] = ye[r] + km[r + NED * c] * pe[c];
  }
  memset(kp, 0, sizeof(double) * ((size_t)neq + 1));
  for (int32_t e = 0; e < nelmt; e++) {
    const int32_t *g = gdof + (size_t)NED * e;
    const double *ye = y + (size_t)NED * e;
    for (int r = 0; r < NED; r++) kp[g[r]] = kp[g[r]] + ye[r];
  }
  kp[0] = 0.0;
  free(y);
}

/* Summation order of the PCG dot products.  0 (default) is the reference's
 * sequential dot_product (src/solver.f90:61-62,72); 1 sums in descending
 * index order and 2 in blocks of 1024 combined pairwise -- both equally
 * valid orders, used only to measure how sensitive the reference's results
 * are to the order of its own sums. */
static int g_dot_mode = 0;
void orc_set_dot_mode(int mode) { g_dot_mode = mode; }

static double dot_tree(const double *a, const double *b, int64_t lo,
                       int64_t hi) {
  if (hi - lo <= 1024) {
    double s = 0.0;
    for (int64_t i = lo; i < hi; i++) s = s + a[i] * b[i];
    return s;
  }
  int64_t mid = lo + (((hi - lo) / 2 + 1023) / 1024) * 1024;
  return dot_tree(a, b, lo, mid) + dot_tree(a, b, mid, hi);
}

static double dot(const double *a, const double *b, int64_t n1) {
  double s = 0.0;
  if (g_dot_mode == 1) {
    for (int64_t i = n1 - 1; i >= 0; i--) s = s + a[i] * b[i];
    return s;
  }
  if (g_dot_mode == 2) return dot_tree(a, b, 0, n1);
  for (int64_t i = 0; i < n1; i++) s = s + a[i] * b[i];
  return s;
}

static double maxabs(const double *a, int64_t n1) {
  double m = 0.0;
  for (int64_t i = 0; i < n1; i++) {
    double v = fabs(a[i]);
    if (v > m) m = v;
  }
  return m;
}

int orc_pcg_solver(int32_t neq, int32_t nelmt, const double *k, double *u,
                   const double *f, const double *dinv, const int32_t *gdof,
                   double cg_tol, int32_t cg_maxiter, int32_t *cg_iter) {
  const int64_t n1 = (int64_t)neq + 1;
  double *kp = (double *)calloc(n1, sizeof(double));
  double *p = (double *)malloc(sizeof(double) * n1);
  double *r = (double *)malloc(sizeof(double) * n1);
  double *z = (double *)malloc(sizeof(double) * n1);
  int errcode = -1;
  if (maxabs(u, n1) > 0.0) orc_ebe_matvec(neq, nelmt, k, gdof, u, kp);
  for (int64_t i = 0; i < n1; i++) r[i] = f[i] - kp[i];
  for (int64_t i = 0; i < n1; i++) z[i] = dinv[i] * r[i];
  memcpy(p, z, sizeof(double) * n1);
  int32_t it;
  for (it = 1; it <= cg_maxiter; it++) {
    orc_ebe_matvec(neq, nelmt, k, gdof, p, kp);
    double rz = dot(r, z, n1);
    double alpha = rz / dot(p, kp, n1);
    for (int64_t i = 0; i < n1; i++) u[i] = u[i] + alpha * p[i];
    if (fabs(alpha) * maxabs(p, n1) / maxabs(u, n1) <= cg_tol) {
      errcode = 0;
      break;
    }
    for (int64_t i = 0; i < n1; i++) r[i] = r[i] - alpha * kp[i];
    for (int64_t i = 0; i < n1; i++) z[i] = dinv[i] * r[i];
    double beta = dot(r, z, n1) / rz;
    for (int64_t i = 0; i < n1; i++) p[i] = z[i] + beta * p[i];
  }
  *cg_iter = it;
  free(kp); free(p); free(r); free(z);
  return errcode;
}

/* ------------------------------------------------------------------------ */
/* Mohr-Coulomb / viscoplastic library                                      */
/* ------------------------------------------------------------------------ */

void orc_stress_invariant(const double *s, double *sigm, double *dsbar,
                          double *theta) {
  /* src/math_library.f90:160-184 */
  const double small = 1.e-12;
  double sq3 = sqrt(3.0);
  double s1 = s[0], s2 = s[1], s3 = s[2], s4 = s[3], s5 = s[4], s6 = s[5];
  *sigm = (s1 + s2 + s3) / 3.0;
  double d2 = ((s1 - s2) * (s1 - s2) + (s2 - s3) * (s2 - s3) +
               (s3 - s1) * (s3 - s1)) / 6.0 + s4 * s4 + s5 * s5 + s6 * s6;
  if (d2 < small) d2 = small;
  double ds1 = s1 - *sigm, ds2 = s2 - *sigm, ds3 = s3 - *sigm;
  double d3 = ds1 * ds2 * ds3 - ds1 * s5 * s5 - ds2 * s6 * s6 -
              ds3 * s4 * s4 + 2.0 * s4 * s5 * s6;
  *dsbar = sq3 * sqrt(d2);
  if (*dsbar < small) {
    *theta = 0.0;
  } else {
    double sd = sqrt(d2);
    double sine = -3.0 * sq3 * d3 / (2.0 * (sd * sd * sd));
    if (sine >= 1.0) sine = 1.0;
    if (sine < -1.0) sine = -1.0;
    *theta = asin(sine) / 3.0;
  }
}

double orc_mohcouf(double phi, double c, double sigm, double dsbar,
                   double theta) {
  /* src/plastic_library.f90:101-106 */
  /* the reference build evaluates each sin/cos pair with one sincos call
   * (nm: plastic_library.o imports sincos); glibc's sincos and separate
   * sin/cos calls differ in the last bit for rare arguments */
  double phir = phi * DEG2RAD;
  double snph, csph, snth, csth;
  sincos(phir, &snph, &csph);
  sincos(theta, &snth, &csth);
  return snph * sigm + dsbar * (csth / sqrt(3.0) - snth * snph / 3.0) -
         c * csph;
}

void orc_mohcouq(double psi, double dsbar, double theta, double *dq1,
                 double *dq2, double *dq3) {
  /* src/plastic_library.f90:123-141 */
  double psir = psi * DEG2RAD;
  double snth = sin(theta), snps = sin(psir), sq3 = sqrt(3.0);
  *dq1 = snps;
  if (fabs(snth) > 0.49) {
    double c1 = 1.0;
    if (snth < 0.0) c1 = -1.0;
    *dq2 = (sq3 * 0.5 - c1 * snps * 0.5 / sq3) * sq3 * 0.5 / dsbar;
    *dq3 = 0.0;
  } else {
    double csth = cos(theta);
    double cs3th = cos(3.0 * theta);
    double tn3th = tan(3.0 * theta);
    double tnth = snth / csth;
    *dq2 = sq3 * csth / dsbar * ((1.0 + tnth * tn3th) + snps * (tn3th - tnth) / sq3) * 0.5;
    *dq3 = 0.5 * 3.0 * (sq3 * snth + snps * csth) / (cs3th * dsbar * dsbar);
  }
}

void orc_formm(const double *s, double *m1, double *m2, double *m3) {
  /* src/plastic_library.f90:164-206; (i,j) at [i + 6*j] */
  double sx = s[0], sy = s[1], sz = s[2], txy = s[3], tyz = s[4], tzx = s[5];
  double sigm = (sx + sy + sz) / 3.0;
  double dx = sx - sigm, dy = sy - sigm, dz = sz - sigm;
  memset(m1, 0, sizeof(double) * 36);
  memset(m2, 0, sizeof(double) * 36);
  memset(m3, 0, sizeof(double) * 36);
#define M(m, i, j) m[((i) - 1) + 6 * ((j) - 1)]
  for (int i = 1; i <= 3; i++)
    for (int j = 1; j <= 3; j++) M(m1, i, j) = 1.0 / (3.0 * sigm);
  for (int i = 1; i <= 3; i++) {
    M(m2, i, i) = 2.0;
    M(m2, i + 3, i + 3) = 6.0;
  }
  M(m2, 1, 2) = -1.0;
  M(m2, 1, 3) = -1.0;
  M(m2, 2, 3) = -1.0;
  M(m3, 1, 1) = dx;
  M(m3, 1, 2) = dz;
  M(m3, 1, 3) = dy;
  M(m3, 1, 4) = txy;
  M(m3, 1, 5) = -2.0 * tyz;
  M(m3, 1, 6) = tzx;
  M(m3, 2, 2) = dy;
  M(m3, 2, 3) = dx;
  M(m3, 2, 4) = txy;
  M(m3, 2, 5) = tyz;
  M(m3, 2, 6) = -2.0 * tzx;
  M(m3, 3, 3) = dz;
  M(m3, 3, 4) = -2.0 * txy;
  M(m3, 3, 5) = tyz;
  M(m3, 3, 6) = tzx;
  M(m3, 4, 4) = -3.0 * dz;
  M(m3, 4, 5) = 3.0 * tzx;
  M(m3, 4, 6) = 3.0 * tyz;
  M(m3, 5, 5) = -3.0 * dx;
  M(m3, 5, 6) = 3.0 * txy;
  M(m3, 6, 6) = -3.0 * dy;
  for (int i = 1; i <= 6; i++)
    for (int j = i + 1; j <= 6; j++) {
      M(m1, j, i) = M(m1, i, j);
      M(m2, j, i) = M(m2, i, j);
      M(m3, j, i) = M(m3, i, j);
    }
#undef M
  for (int q = 0; q < 36; q++) {
    m1[q] = m1[q] / 3.0;
    m2[q] = m2[q] / 3.0;
    m3[q] = m3[q] / 3.0;
  }
}

double orc_dt_viscoplas(int nmat, const double *nuf, const double *phif,
                        const double *ymf) {
  /* src/plastic_library.f90:31-38 */
  double dt_min = 1.0e32;
  for (int m = 0; m < nmat; m++) {
    double snphi = sin(phif[m] * DEG2RAD);
    double dt = 4.0 * (1.0 + nuf[m]) * (1.0 - 2.0 * nuf[m]) /
                (ymf[m] * (1.0 - 2.0 * nuf[m] + snphi * snphi));
    if (dt < dt_min) dt_min = dt;
  }
  return dt_min;
}

int orc_strength_reduction(double srf, int phinu, int nmat, const double *coh,
                           const double *nu, const double *phi,
                           const double *psi, double *cohf, double *nuf,
                           double *phif, double *psif) {
  /* src/plastic_library.f90:58-87 */
  for (int m = 0; m < nmat; m++) {
    cohf[m] = coh[m]; nuf[m] = nu[m]; phif[m] = phi[m]; psif[m] = psi[m];
  }
  int istat = 0;
  if (srf != 1.0) {
    for (int m = 0; m < nmat; m++) {
      double tnphi = tan(phi[m] * DEG2RAD);
      phif[m] = atan(tnphi / srf) * RAD2DEG;
      double tnpsi = tan(psi[m] * DEG2RAD);
      psif[m] = atan(tnpsi / srf) * RAD2DEG;
      cohf[m] = coh[m] / srf;
    }
  }
  if (!phinu) return istat;
  for (int m = 0; m < nmat; m++) {
    double omtnu = 1.0 - 2.0 * nu[m];
    double snphif = sin(phif[m] * DEG2RAD);
    if (snphif < omtnu) {
      double snphi = sin(phi[m] * DEG2RAD);
      double beta_nuphi = snphi / omtnu;
      if (beta_nuphi < 1.0) beta_nuphi = 1.0 + 1.0e-12;
      nuf[m] = 0.5 * (1.0 - snphif / beta_nuphi);
      istat = 1;
    }
  }
  return istat;
}

/* ------------------------------------------------------------------------ */
/* plastic-iteration body                                                   */
/* ------------------------------------------------------------------------ */

/* One element (src/slope.F90:316-378 / src/excavation.F90:550-612).
 * excav: effective stress = (sigma [+ water]) + stress_elmt, else
 * sigma + stress_elmt; identical values here (no water) -- the order only
 * matters with water.  Returns fmax over the element's points. */
static double body_element(const tables_t *T, const double *xc,
                           const int32_t *g, double E, double nuf, double phif,
                           double cohf, double psif, double dt, const double *x,
                           double *evpt, double *stress, int conv_or_last,
                           double *bload, uint8_t *flags, double *vm_pt,
                           double *sf_pt) {
  double cmat[36], deriv[3 * NG], b[NST * NED], eld[NED];
  orc_compute_cmat(E, nuf, cmat);
  for (int j = 0; j < NED; j++) eld[j] = x[g[j]];
  for (int j = 0; j < NED; j++) bload[j] = 0.0;
  double fmax = 0.0;
  for (int i = 0; i < NG; i++) {
    double detjac = gll_point_deriv(T, xc, i, deriv);
    compute_bmat(deriv, b);
    double eps[NST], sigma[NST], eff[NST];
    for (int s = 0; s < NST; s++) {
      double acc = 0.0;
      for (int j = 0; j < NED; j++) acc = acc + b[s + 6 * j] * eld[j];
      eps[s] = acc;
    }
    double *ev = evpt + 6 * i;
    for (int s = 0; s < NST; s++) eps[s] = eps[s] - ev[s];
    for (int s = 0; s < NST; s++) {
      double acc = 0.0;
      for (int t = 0; t < NST; t++) acc = acc + cmat[s + 6 * t] * eps[t];
      sigma[s] = acc;
    }
    double *st = stress + 6 * i;
    for (int s = 0; s < NST; s++) eff[s] = sigma[s] + st[s];
    double sigm, dsbar, theta;
    orc_stress_invariant(eff, &sigm, &dsbar, &theta);
    double f = orc_mohcouf(phif, cohf, sigm, dsbar, theta);
    if (f > fmax) fmax = f;
    if (flags) flags[i] = (f >= 0.0);
    if (f >= 0.0) {
      double dq1, dq2, dq3, m1[36], m2[36], m3[36], flow[36];
      orc_mohcouq(psif, dsbar, theta, &dq1, &dq2, &dq3);
      orc_formm(eff, m1, m2, m3);
      for (int q = 0; q < 36; q++)
        flow[q] = f * (m1[q] * dq1 + m2[q] * dq2 + m3[q] * dq3);
      double erate[NST], evp[NST], devp[NST];
      for (int a = 0; a < NST; a++) {
        double acc = 0.0;
        for (int c = 0; c < NST; c++) acc = acc + flow[a + 6 * c] * eff[c];
        erate[a] = acc;
      }
      for (int a = 0; a < NST; a++) evp[a] = erate[a] * dt;
      for (int a = 0; a < NST; a++) ev[a] = ev[a] + evp[a];
      for (int a = 0; a < NST; a++) {
        double acc = 0.0;
        for (int c = 0; c < NST; c++) acc = acc + cmat[a + 6 * c] * evp[c];
        devp[a] = acc;
      }
      if (!conv_or_last) {
        for (int j = 0; j < NED; j++) {
          double acc = 0.0; /* eload = matmul(devp, bmat) */
          for (int s = 0; s < NST; s++) acc = acc + devp[s] * b[s + 6 * j];
          bload[j] = bload[j] + acc * detjac * T->w[i];
        }
      }
    }
    if (conv_or_last) {
      if (vm_pt) { /* src/excavation.F90:601-602 */
        double d = 0.0;
        for (int s = 0; s < NST; s++) d = d + ev[s] * ev[s];
        vm_pt[i] = sqrt((2.0 / 3.0) * d);
      }
      if (sf_pt) { /* src/excavation.F90:604-608 */
        const double phir = phif * DEG2RAD;
        sf_pt[i] = (sigm * sin(phir) - cohf * cos(phir)) /
                   (-dsbar * (cos(theta) / sqrt(3.0) - sin(theta) * sin(phir) / 3.0));
      }
      for (int s = 0; s < NST; s++) st[s] = eff[s];
    }
  }
  return fmax;
}

double orc_plastic_body(int32_t neq, int32_t nelmt, const double *xc,
                        const int32_t *gdof, const int32_t *mat_id,
                        const double *ym, const double *nuf,
                        const double *phif, const double *cohf,
                        const double *psif, double dt, const double *x,
                        double *evpt, double *stress_elmt, double *bodyload,
                        int conv_or_last, uint8_t *flags) {
  (void)neq;
  const tables_t *T = tables();
  double *bl = (double *)malloc(sizeof(double) * NED * (size_t)nelmt);
  double fmax = 0.0;
#pragma omp parallel for schedule(static) reduction(max : fmax)
  for (int32_t e = 0; e < nelmt; e++) {
    int m = mat_id[e] - 1;
    double fe = body_element(T, xc + 24 * (size_t)e, gdof + (size_t)NED * e,
                             ym[m], nuf[m], phif[m], cohf[m], psif[m], dt, x,
                             evpt + 6 * NG * (size_t)e,
                             stress_elmt + 6 * NG * (size_t)e, conv_or_last,
                             bl + (size_t)NED * e,
                             flags ? flags + NG * (size_t)e : NULL, NULL, NULL);
    if (fe > fmax) fmax = fe;
  }
  if (!conv_or_last) {
    for (int32_t e = 0; e < nelmt; e++) {
      const int32_t *g = gdof + (size_t)NED * e;
      const double *b = bl + (size_t)NED * e;
      double tmp[NED];
      for (int j = 0; j < NED; j++) tmp[j] = bodyload[g[j]] + b[j];
      for (int j = 0; j < NED; j++) bodyload[g[j]] = tmp[j];
    }
  }
  bodyload[0] = 0.0;
  free(bl);
  return fmax;
}

double orc_plastic_body_excav(int32_t neq, int32_t nel_act,
                              const int32_t *elmt, const double *xc,
                              const int32_t *gdof_act, const int32_t *mat_id,
                              const double *ym, const double *nuf,
                              const double *phif, const double *cohf,
                              const double *psif, double dt, const double *x,
                              double *evpt, double *stress_elmt,
                              double *bodyload, int conv_or_last,
                              const int32_t *g_num, double *vmeps, double *scf) {
  (void)neq;
  const tables_t *T = tables();
  double *bl = (double *)malloc(sizeof(double) * NED * (size_t)nel_act);
  double *vm = (double *)malloc(sizeof(double) * NG * (size_t)nel_act);
  double *sf = (double *)malloc(sizeof(double) * NG * (size_t)nel_act);
  double fmax = 0.0;
#pragma omp parallel for schedule(static) reduction(max : fmax)
  for (int32_t a = 0; a < nel_act; a++) {
    int32_t e = elmt[a];
    int m = mat_id[e] - 1;
    double fe = body_element(T, xc + 24 * (size_t)e, gdof_act + (size_t)NED * a,
                             ym[m], nuf[m], phif[m], cohf[m], psif[m], dt, x,
                             evpt + 6 * NG * (size_t)e,
                             stress_elmt + 6 * NG * (size_t)e, conv_or_last,
                             bl + (size_t)NED * a, NULL, vm + NG * (size_t)a,
                             sf + NG * (size_t)a);
    if (fe > fmax) fmax = fe;
  }
  if (!conv_or_last) {
    for (int32_t a = 0; a < nel_act; a++) {
      const int32_t *g = gdof_act + (size_t)NED * a;
      const double *b = bl + (size_t)NED * a;
      double tmp[NED];
      for (int j = 0; j < NED; j++) tmp[j] = bodyload[g[j]] + b[j];
      for (int j = 0; j < NED; j++) bodyload[g[j]] = tmp[j];
    }
  } else if (g_num) { /* element then point order (src/excavation.F90:546-609) */
    for (int32_t a = 0; a < nel_act; a++) {
      int32_t e = elmt[a];
      for (int i = 0; i < NG; i++) {
        int32_t n = g_num[NG * (size_t)e + i];
        if (vmeps) vmeps[n] = vmeps[n] + vm[NG * (size_t)a + i];
        if (scf && sf[NG * (size_t)a + i] < scf[n]) scf[n] = sf[NG * (size_t)a + i];
      }
    }
  }
  bodyload[0] = 0.0;
  free(bl);
  free(vm);
  free(sf);
  return fmax;
}

/* ------------------------------------------------------------------------ */
/* slope3d SRF sweep (src/slope.F90:244-443)                                */
/* ------------------------------------------------------------------------ */

int orc_slope3d(int32_t nelmt, int32_t nnode, int32_t neq, const double *xc,
                const int32_t *gdof, const int32_t *gdof_node,
                const int32_t *mat_id, int nmat, const double *gam,
                const double *ym, const double *nu, const double *phi,
                const double *coh, const double *psi, int phinu,
                double cg_tol, int32_t cg_maxiter, double nl_tol,
                int32_t nl_maxiter, int nsrf, const double *srf,
                int32_t *sum_cg, int32_t *sum_nl, double *uxmax, double *umax,
                double *fmax_out, double *x_out, int32_t *cg_log,
                int32_t cg_log_cap, int32_t *n_cg_log, double *stress_out) {
  const int64_t n1 = (int64_t)neq + 1;
  const size_t nk = (size_t)NED * NED * nelmt;
  double *storekm = (double *)malloc(sizeof(double) * nk);
  double *dprecon = (double *)calloc(n1, sizeof(double));
  double *extload = (double *)calloc(n1, sizeof(double));
  double *bodyload = (double *)calloc(n1, sizeof(double));
  double *load = (double *)calloc(n1, sizeof(double));
  double *x = (double *)calloc(n1, sizeof(double));
  double *oldx = (double *)calloc(n1, sizeof(double));
  double *evpt = (double *)calloc((size_t)6 * NG * nelmt, sizeof(double));
  double *stress = (double *)calloc((size_t)6 * NG * nelmt, sizeof(double));
  double *cohf = (double *)malloc(sizeof(double) * nmat);
  double *nuf = (double *)malloc(sizeof(double) * nmat);
  double *phif = (double *)malloc(sizeof(double) * nmat);
  double *psif = (double *)malloc(sizeof(double) * nmat);
  int nrun = 0, ncg = 0, rc = 0;

  orc_stiffness_bodyload(nelmt, neq, xc, gdof, mat_id, gam, nu, ym, storekm,
                         dprecon, extload, NULL);
  for (int64_t i = 1; i < n1; i++) dprecon[i] = 1.0 / dprecon[i];
  dprecon[0] = 0.0;

  for (int s = 0; s < nsrf; s++) {
    memset(stress, 0, sizeof(double) * 6 * NG * (size_t)nelmt);
    int istat = orc_strength_reduction(srf[s], phinu, nmat, coh, nu, phi, psi,
                                       cohf, nuf, phif, psif);
    double dt = orc_dt_viscoplas(nmat, nuf, phif, ym);
    if (istat == 1) {
      orc_stiffness_bodyload(nelmt, neq, xc, gdof, mat_id, gam, nuf, ym,
                             storekm, dprecon, NULL, NULL);
      for (int64_t i = 1; i < n1; i++) dprecon[i] = 1.0 / dprecon[i];
      dprecon[0] = 0.0;
    }
    int32_t cg_tot = 0, nl_iter;
    memset(bodyload, 0, sizeof(double) * n1);
    memset(evpt, 0, sizeof(double) * 6 * NG * (size_t)nelmt);
    memset(x, 0, sizeof(double) * n1);
    memset(oldx, 0, sizeof(double) * n1);
    double fmax = 0.0;
    int conv = 0;
    for (nl_iter = 1; nl_iter <= nl_maxiter; nl_iter++) {
      fmax = 0.0;
      for (int64_t i = 0; i < n1; i++) load[i] = extload[i] + bodyload[i];
      load[0] = 0.0;
      int32_t cg_iter = 0;
      int ec = orc_pcg_solver(neq, nelmt, storekm, x, load, dprecon, gdof,
                              cg_tol, cg_maxiter, &cg_iter);
      if (cg_log && ncg < cg_log_cap) cg_log[ncg] = cg_iter;
      ncg++;
      if (ec != 0) { rc = -1 - s; goto done; }
      cg_tot += cg_iter;
      x[0] = 0.0;
      double num = 0.0;
      for (int64_t i = 0; i < n1; i++) {
        double d = fabs(x[i] - oldx[i]);
        if (d > num) num = d;
      }
      double uerr = num / maxabs(x, n1);
      memcpy(oldx, x, sizeof(double) * n1);
      conv = uerr <= nl_tol;
      int col = conv || nl_iter == nl_maxiter;
      fmax = orc_plastic_body(neq, nelmt, xc, gdof, mat_id, ym, nuf, phif,
                              cohf, psif, dt, x, evpt, stress, bodyload, col,
                              NULL);
      if (col) break;
    }
    if (nl_iter > nl_maxiter) nl_iter = nl_maxiter;
    sum_cg[s] = cg_tot;
    sum_nl[s] = nl_iter;
    /* nodalu / uxmax / umax (src/slope.F90:401-413) */
    double ux = 0.0, um = 0.0;
    for (int32_t n = 0; n < nnode; n++) {
      double u3[3];
      for (int c = 0; c < 3; c++) {
        int32_t q = gdof_node[c + 3 * (size_t)n];
        u3[c] = q ? x[q] : 0.0;
        if (fabs(u3[c]) > ux) ux = fabs(u3[c]);
      }
      double mag = sqrt(u3[0] * u3[0] + u3[1] * u3[1] + u3[2] * u3[2]);
      if (mag > um) um = mag;
    }
    uxmax[s] = ux;
    umax[s] = um;
    fmax_out[s] = fmax;
    nrun = s + 1;
    if (x_out) memcpy(x_out, x, sizeof(double) * n1);
    if (stress_out) memcpy(stress_out, stress, sizeof(double) * 6 * NG * (size_t)nelmt);
    if (nl_iter == nl_maxiter) break;
  }
done:
  if (n_cg_log) *n_cg_log = ncg;
  free(storekm); free(dprecon); free(extload); free(bodyload); free(load);
  free(x); free(oldx); free(evpt); free(stress); free(cohf); free(nuf);
  free(phif); free(psif);
  return rc < 0 ? rc : nrun;
}

/* ------------------------------------------------------------------------ */
/* nodal result fields (src/slope.F90:203-210, 418-437;                     */
/*                      src/postprocess.f90:146-183)                        */
/* ------------------------------------------------------------------------ */

void orc_nodal_fields(int32_t nelmt, int32_t nnode, const int32_t *g_num,
                      const double *stress, double *sig, double *psig,
                      double *nsig) {
  int32_t *val = (int32_t *)calloc((size_t)nnode, sizeof(int32_t));
  /* node_valency (src/slope.F90:205-210) */
  for (int64_t q = 0; q < (int64_t)nelmt * NG; q++) val[g_num[q]]++;
  /* stress_global: element order, then the mean (src/slope.F90:424-437) */
  memset(sig, 0, sizeof(double) * NST * (size_t)nnode);
  for (int64_t e = 0; e < nelmt; e++)
    for (int i = 0; i < NG; i++) {
      const int64_t n = g_num[e * NG + i];
      for (int c = 0; c < NST; c++)
        sig[n * NST + c] = sig[n * NST + c] + stress[(e * NG + i) * NST + c];
    }
  for (int64_t n = 0; n < nnode; n++)
    for (int c = 0; c < NST; c++) sig[n * NST + c] = sig[n * NST + c] / (double)val[n];
  /* principal stresses, ascending, and the normal stress on the plane of
   * maximum shear (src/postprocess.f90:146-183, rquick_sort
   * src/math_library.f90:211-228) */
  const double two_third = 2.0 / 3.0;
  for (int64_t n = 0; n < nnode; n++) {
    double sigm, dsbar, theta;
    orc_stress_invariant(sig + n * NST, &sigm, &dsbar, &theta);
    const double a = two_third * dsbar;
    double p[3] = {sigm + a * sin(theta - two_third * PI_REF), sigm + a * sin(theta),
                   sigm + a * sin(theta + two_third * PI_REF)};
    for (int i = 1; i < 3; i++) {
      const double t = p[i];
      int j = i - 1;
      while (j >= 0 && p[j] > t) {
        p[j + 1] = p[j];
        j--;
      }
      p[j + 1] = t;
    }
    if (psig)
      for (int k = 0; k < 3; k++) psig[n * 3 + k] = p[k];
    if (nsig) nsig[n] = 0.5 * (p[0] + p[2]);
  }
  free(val);
}

/* ------------------------------------------------------------------------ */
/* overburden initial stress (src/postprocess.f90:10-33), s0_type 1         */
/* ------------------------------------------------------------------------ */

void orc_overburden_stress(int32_t nelmt, const int32_t *g_num, const int32_t *mat_id,
                           const double *node_z, const double *gam, double z0,
                           double p0, double k0, double *stress) {
  /* stress_elmt starts at zero (src/excavation.F90:180); only the normal
   * components are set */
  memset(stress, 0, sizeof(double) * 6 * NG * (size_t)nelmt);
  for (int64_t e = 0; e < nelmt; e++)
    for (int i = 0; i < NG; i++) {
      const double z = node_z[g_num[e * NG + i]];
      const double szz = p0 - gam[mat_id[e] - 1] * fabs(z - z0); /* compression - */
      double *st = stress + (e * NG + i) * 6;
      st[2] = szz;
      st[0] = k0 * szz;
      st[1] = k0 * szz;
    }
}

/* ------------------------------------------------------------------------ */
/* excavation helpers (src/postprocess.f90:40-79;                           */
/*                     src/excavation_library.f90:186-236)                  */
/* ------------------------------------------------------------------------ */

void orc_elastic_stress(int32_t nel, const int32_t *elmt, const double *xc,
                        const int32_t *gdof, const int32_t *mat_id,
                        const double *ym, const double *nu, const double *x,
                        double *stress) {
  const tables_t *T = tables();
#pragma omp parallel for schedule(static)
  for (int32_t a = 0; a < nel; a++) {
    const int32_t e = elmt ? elmt[a] : a;
    const int m = mat_id[e] - 1;
    double cmat[36], deriv[3 * NG], b[NST * NED], eld[NED];
    orc_compute_cmat(ym[m], nu[m], cmat);
    const int32_t *g = gdof + (size_t)NED * e;
    for (int j = 0; j < NED; j++) eld[j] = x[g[j]];
    for (int i = 0; i < NG; i++) {
      gll_point_deriv(T, xc + 24 * (size_t)e, i, deriv);
      compute_bmat(deriv, b);
      double eps[NST];
      for (int s = 0; s < NST; s++) { /* eps = matmul(bmat, eld) */
        double acc = 0.0;
        for (int j = 0; j < NED; j++) acc = acc + b[s + 6 * j] * eld[j];
        eps[s] = acc;
      }
      double *st = stress + 6 * (i + NG * (size_t)e);
      for (int s = 0; s < NST; s++) { /* sigma = matmul(cmat, eps) */
        double acc = 0.0;
        for (int t = 0; t < NST; t++) acc = acc + cmat[s + 6 * t] * eps[t];
        st[s] = acc;
      }
    }
  }
}

void orc_excavation_load_nodal(int32_t nvoid, const int32_t *elmt_void,
                               const double *xc, const int32_t *g_num,
                               const int32_t *mat_id, const double *gam,
                               const double *stress, int32_t nnode,
                               double *excavload) {
  const tables_t *T = tables();
  memset(excavload, 0, sizeof(double) * 3 * (size_t)nnode);
  for (int32_t a = 0; a < nvoid; a++) {
    const int32_t e = elmt_void[a];
    double deriv[3 * NG], b[NST * NED], bload[NED], eld[NED], tload[NED];
    for (int j = 0; j < NED; j++) { bload[j] = 0.0; eld[j] = 0.0; }
    for (int i = 0; i < NG; i++) {
      double detjac = gll_point_deriv(T, xc + 24 * (size_t)e, i, deriv);
      compute_bmat(deriv, b);
      const double *sg = stress + 6 * (i + NG * (size_t)e);
      for (int j = 0; j < NED; j++) { /* eload = matmul(sigma, bmat) */
        double acc = 0.0;
        for (int s = 0; s < NST; s++) acc = acc + sg[s] * b[s + 6 * j];
        bload[j] = bload[j] + acc * detjac * T->w[i];
      }
      eld[3 * i + 2] = eld[3 * i + 2] + detjac * T->w[i];
    }
    const double gm = gam[mat_id[e] - 1];
    for (int j = 0; j < NED; j++) tload[j] = eld[j] * gm + bload[j];
    const int32_t *num = g_num + (size_t)NG * e;
    for (int c = 0; c < 3; c++)
      for (int n = 0; n < NG; n++)
        excavload[3 * (size_t)num[n] + c] = excavload[3 * (size_t)num[n] + c] +
                                            tload[3 * n + c];
  }
}

double orc_dt_viscoplas_mask(int nmat, const double *nuf, const double *phif,
                             const double *ymf, const uint8_t *ismat) {
  /* src/plastic_library.f90:17-39 with the optional ismat mask */
  double dt_min = 1.0e32;
  for (int m = 0; m < nmat; m++) {
    if (ismat && !ismat[m]) continue;
    double snphi = sin(phif[m] * DEG2RAD);
    double dt = 4.0 * (1.0 + nuf[m]) * (1.0 - 2.0 * nuf[m]) /
                (ymf[m] * (1.0 - 2.0 * nuf[m] + snphi * snphi));
    if (dt < dt_min) dt_min = dt;
  }
  return dt_min;
}
