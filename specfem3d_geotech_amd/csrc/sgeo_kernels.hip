// sgeo_kernels.hip -- CDNA4 (gfx950) element kernels of the SPECFEM3D_GEOTECH
// hot path: element stiffness (src/preprocess.f90:74-148) and the
// viscoplastic Mohr-Coulomb plastic-iteration body (src/slope.F90:312-383,
// src/excavation.F90:546-616).
//
// Floating-point contraction is disabled in this file: every product and sum
// is rounded as written, in the order the reference evaluates it, so the
// element-level arithmetic matches the reference bit for bit; only device
// transcendentals (sin, cos, tan, asin) can differ from the C library's in
// the last bit.
#pragma clang fp contract(off)

#include "sgeo_device.h"
#include "sgeo_element.h"
#include "sgeo_internal.h"

namespace sgeo {

// --------------------------------------------------------------------------
// element stiffness: src/preprocess.f90:108-131
// --------------------------------------------------------------------------
// One block of 256 threads per element.  Thread t owns entries
// idx = t + 256*q of K_e (column-major, r = idx % 81, c = idx / 81).
// Per GLL point: T = B^T C (81x6) in LDS, then K += (T B)*detjac*w, with
// both products summed over their full inner index in ascending order as
// matmul(matmul(transpose(bmat),cmat),bmat) does (zero terms are exact).
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
      // nonzero B(t, r) of column r (src/preprocess.f90:56-64)
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
    const double w = gll_weight(i);
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
    // eld(3i) = detjac*w (src/preprocess.f90:123-125)
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
// element x GLL plastic body (src/slope.F90:312-383; the excavation variant
// src/excavation.F90:546-616 when excav != 0), for one or many trials
// --------------------------------------------------------------------------
// grid.x: element blocks of BODY_EPB elements; grid.y: trial columns
// (cols[blockIdx.y], or column 0).  Trial column `col` reads x[g*xs + col],
// its state at offset col*nel (evpt/stress/BL/flags), its materials at
// mp + col*nmat, its scalars at ps[col].
#define BODY_EPB 9
#define BODY_THREADS 256

struct PointOut {
  double jinv[9];
  double det;
  double devp[6];
};

__global__ __launch_bounds__(BODY_THREADS) void k_plastic_body(
    int32_t nel, const int32_t* __restrict__ elmt,
    const double* __restrict__ corners, const int32_t* __restrict__ gdof,
    const int32_t* __restrict__ mat_id, const MatParams* __restrict__ mp,
    int nmat, const double* __restrict__ x, int xs,
    double* __restrict__ evpt, double* __restrict__ stress,
    double* __restrict__ BL, uint8_t* __restrict__ flags,
    PlasticScalars* __restrict__ ps, int is_last,
    const int32_t* __restrict__ is_last_col, const int32_t* __restrict__ cols,
    int excav, double* __restrict__ vm_pt, double* __restrict__ sf_pt) {
  __shared__ double s_x[BODY_EPB][81];
  __shared__ double s_xc[BODY_EPB][24];
  __shared__ PointOut s_pt[BODY_EPB][27];
  __shared__ uint8_t s_flag[BODY_EPB][27];
  __shared__ double red[BODY_THREADS / 64];
  const int col = cols ? cols[blockIdx.y] : 0;
  PlasticScalars* pc = ps + col;
  const int last = is_last_col ? is_last_col[col] : is_last;
  const int conv_or_last = pc->nl_isconv || last;
  const int64_t so = (int64_t)col * nel;  // state offset in elements
  const MatParams* mpc = mp + (int64_t)col * nmat;
  const int t = threadIdx.x;
  const int64_t e0 = (int64_t)blockIdx.x * BODY_EPB;
  for (int q = t; q < BODY_EPB * 81; q += BODY_THREADS) {
    const int el = q / 81, j = q % 81;
    const int64_t a = e0 + el;
    if (a < nel) s_x[el][j] = x[(int64_t)gdof[a * 81 + j] * xs + col];
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
    const MatParams mm = mpc[mat_id[e] - 1];
    double jinv[3][3];
    const double det = jacobian_point(i, s_xc[el], jinv);
    double* ev = evpt + ((so + a) * 27 + i) * 6;
    double* st = stress + ((so + a) * 27 + i) * 6;
    double evl[6], stl[6], eff[6], devp[6];
    for (int s = 0; s < 6; s++) { evl[s] = ev[s]; stl[s] = st[s]; }
    double inv[3];
    const double f = point_update(i, jinv, s_x[el], mm, stl, evl, eff, devp, inv);
    fmax = f > 0.0 ? f : 0.0;  // fmax starts at zero (src/slope.F90:286)
    const bool plastic = f >= 0.0;
    if (flags) flags[(so + a) * 27 + i] = plastic ? 1 : 0;
    nplast = plastic ? 1 : 0;
    if (plastic)
      for (int s = 0; s < 6; s++) ev[s] = evl[s];
    const bool contrib = plastic && !conv_or_last;
    s_flag[el][i] = contrib ? 1 : 0;
    PointOut& po = s_pt[el][i];
    po.det = det;
    for (int p = 0; p < 3; p++)
      for (int q = 0; q < 3; q++) po.jinv[p * 3 + q] = jinv[p][q];
    for (int p = 0; p < 6; p++) po.devp[p] = devp[p];
    if (conv_or_last) {
      if (excav && vm_pt) {  // src/excavation.F90:601-602
        double d = 0.0;
        for (int s = 0; s < 6; s++) d = d + evl[s] * evl[s];
        vm_pt[(so + a) * 27 + i] = sqrt((2.0 / 3.0) * d);
      }
      if (excav && sf_pt) {  // stress concentration factor, src/excavation.F90:604-608
        const double phir = mm.phif * (3.141592653589793 / 180.0);
        const double snph = sin(phir);
        sf_pt[(so + a) * 27 + i] =
            (inv[0] * snph - mm.cohf * cos(phir)) /
            (-inv[1] * (cos(inv[2]) / sqrt(3.0) - sin(inv[2]) * snph / 3.0));
      }
      for (int s = 0; s < 6; s++) st[s] = eff[s];  // stress_elmt = effsigma
    }
  }
  // max f and plastic count (order independent)
  const double bmax = block_max<BODY_THREADS / 64>(fmax, red);
  const double bcnt = block_sum<BODY_THREADS / 64>((double)nplast, red);
  if (t == 0) {
    atomicMax((unsigned long long*)&pc->fmax_bits,
              (unsigned long long)__double_as_longlong(bmax));
    atomicAdd(&pc->n_plastic, (int)bcnt);
  }
  if (conv_or_last) return;  // no body load (src/slope.F90:380)
  __syncthreads();
  // bload(j) = sum_i (eload_i(j)*detjac_i)*w_i over the plastic points,
  // ascending i (src/slope.F90:361-362)
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
      bl = bl + eload_dof(comp, po.devp, d) * po.det * gll_weight(i2);
    }
    BL[(so + a2) * 81 + j] = bl;
  }
}

// dinv = 1/d (index 0 -> 0), src/slope.F90:199
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
// elastic stress at the GLL points (src/postprocess.f90:40-79): thread per
// (element, point); sigma = matmul(cmat, matmul(bmat, u_e)), cmat(E, nu)
// --------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_elastic_stress(
    int32_t nelmt, const double* __restrict__ corners, const int32_t* __restrict__ gdof,
    const int32_t* __restrict__ mat_id, const double* __restrict__ ym,
    const double* __restrict__ nu, const double* __restrict__ x,
    double* __restrict__ stress) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= (int64_t)nelmt * 27) return;
  const int64_t e = q / 27;
  const int i = (int)(q - e * 27);
  double xc[24];
  for (int k = 0; k < 24; k++) xc[k] = corners[e * 24 + k];
  double jinv[3][3];
  jacobian_point(i, xc, jinv);
  const int m = mat_id[e] - 1;
  const Cmat cm = cmat_of(ym[m], nu[m]);
  double eps[6] = {0, 0, 0, 0, 0, 0};
  for (int n = 0; n < 27; n++) {  // eps = matmul(bmat, eld), zero terms skipped
    double d[3];
    deriv_node(i, n, jinv, d);
    const double ux = x[gdof[e * 81 + 3 * n]], uy = x[gdof[e * 81 + 3 * n + 1]],
                 uz = x[gdof[e * 81 + 3 * n + 2]];
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
  cmat_apply(cm, eps, sig);
  for (int s = 0; s < 6; s++) stress[q * 6 + s] = sig[s];
}

// --------------------------------------------------------------------------
// excavation load of the removed elements (src/excavation_library.f90:
// 186-236): Y[e][j] = eld(j)*gam + sum_i eload_i(j) detJ_i w_i, eload_i =
// matmul(sigma_i, bmat_i), eld(3i+3) = detJ_i w_i.  One block per element:
// the 27 points' inverse Jacobians and stresses in LDS, thread j sums its
// dof over the points in order.
// --------------------------------------------------------------------------
__global__ __launch_bounds__(128) void k_excav_load(
    int32_t nelmt, const double* __restrict__ corners,
    const int32_t* __restrict__ mat_id, const double* __restrict__ gam,
    const double* __restrict__ stress, double* __restrict__ Y) {
  __shared__ double s_jinv[27][9];
  __shared__ double s_det[27];
  __shared__ double s_sig[27][6];
  const int64_t e = blockIdx.x;
  const int t = threadIdx.x;
  if (t < 27) {
    double xc[24];
    for (int k = 0; k < 24; k++) xc[k] = corners[e * 24 + k];
    double jinv[3][3];
    s_det[t] = jacobian_point(t, xc, jinv);
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) s_jinv[t][a * 3 + b] = jinv[a][b];
    for (int c = 0; c < 6; c++) s_sig[t][c] = stress[(e * 27 + t) * 6 + c];
  }
  __syncthreads();
  if (t >= 81) return;
  const int n = t / 3, comp = t % 3;
  double bl = 0.0;
  for (int i = 0; i < 27; i++) {
    double jinv[3][3];
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) jinv[a][b] = s_jinv[i][a * 3 + b];
    double d[3];
    deriv_node(i, n, jinv, d);
    bl = bl + eload_dof(comp, s_sig[i], d) * s_det[i] * gll_weight(i);
  }
  const double eld = (comp == 2) ? s_det[n] * gll_weight(n) : 0.0;
  Y[e * 81 + t] = eld * gam[mat_id[e] - 1] + bl;
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

void launch_elastic_stress(hipStream_t s, int32_t nelmt, const double* corners,
                           const int32_t* gdof, const int32_t* mat_id, const double* ym,
                           const double* nu, const double* x, double* stress) {
  if (nelmt <= 0) return;
  hipLaunchKernelGGL(k_elastic_stress, dim3(nblk((int64_t)nelmt * 27, 256)), dim3(256),
                     0, s, nelmt, corners, gdof, mat_id, ym, nu, x, stress);
}

// --------------------------------------------------------------------------
// nodal result fields (src/slope.F90:418-437, src/postprocess.f90:146-183):
// thread per node.  The stresses at the node's points (`slots` = e*27+i,
// ascending in e) are summed from zero in that order and divided by the
// node's valency in FP64; the principal stresses follow from the
// invariants, ascending, and the normal stress on the plane of maximum
// shear is their mean.  Outputs are rounded to FP32, component blocks of
// nnode values, as the EnSight writers store them.  Any output may be null.
// --------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_nodal_fields(
    int32_t nnode, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ slots,
    const double* __restrict__ stress, float* __restrict__ sig, float* __restrict__ psig,
    float* __restrict__ nsig) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (n >= nnode) return;
  const int32_t q0 = rowptr[n], q1 = rowptr[n + 1];
  const double val = (double)(q1 - q0);  // node_valency
  double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int32_t q = q0; q < q1; q++) {
    const double* p = stress + (int64_t)slots[q] * 6;
#pragma unroll
    for (int c = 0; c < 6; c++) s[c] = s[c] + p[c];
  }
#pragma unroll
  for (int c = 0; c < 6; c++) s[c] = s[c] / val;
  if (sig)
#pragma unroll
    for (int c = 0; c < 6; c++) sig[(int64_t)c * nnode + n] = (float)s[c];
  if (!(psig || nsig)) return;
  double sigm, dsbar, theta;
  stress_invariant(s, sigm, dsbar, theta);
  const double two_third = 2.0 / 3.0;
  const double pi = 3.141592653589793;  // src/global.f90:19
  const double a = two_third * dsbar;
  double ps[3] = {sigm + a * sin(theta - two_third * pi), sigm + a * sin(theta),
                  sigm + a * sin(theta + two_third * pi)};
  for (int i = 1; i < 3; i++) {  // insertion sort, ascending (src/math_library.f90:211-228)
    const double t = ps[i];
    int j = i - 1;
    while (j >= 0 && ps[j] > t) {
      ps[j + 1] = ps[j];
      j--;
    }
    ps[j + 1] = t;
  }
  if (psig)
    for (int k = 0; k < 3; k++) psig[(int64_t)k * nnode + n] = (float)ps[k];
  if (nsig) nsig[n] = (float)(0.5 * (ps[0] + ps[2]));
}

// Running nodal sums of the excavation's per-point plastic strain and
// minima of its stress concentration factor over a stage's load increments
// (src/excavation.F90:599-609): per node, the previous value is carried
// through the node's points in element order, as the reference's element
// loop does.  first != 0 starts from 0 and inftol (src/excavation.F90:335).
__global__ __launch_bounds__(256) void k_nodal_accumulate(
    int32_t nnode, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ slots,
    const double* __restrict__ vm, const double* __restrict__ sf,
    double* __restrict__ acc, int first) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (n >= nnode) return;
  double v = first ? 0.0 : acc[n];
  double m = first ? 1.0e32 : acc[nnode + n];
  for (int32_t q = rowptr[n]; q < rowptr[n + 1]; q++) {
    const int32_t p = slots[q];
    v = v + vm[p];
    if (sf[p] < m) m = sf[p];
  }
  acc[n] = v;
  acc[nnode + n] = m;
}

// FP32 outputs of the accumulated fields: vmeps = sum / valency, scf.
__global__ __launch_bounds__(256) void k_nodal_acc_out(int32_t nnode,
                                                       const int32_t* __restrict__ rowptr,
                                                       const double* __restrict__ acc,
                                                       float* __restrict__ vmo,
                                                       float* __restrict__ scfo) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (n >= nnode) return;
  if (vmo) vmo[n] = (float)(acc[n] / (double)(rowptr[n + 1] - rowptr[n]));
  if (scfo) scfo[n] = (float)acc[nnode + n];
}

void launch_nodal_accumulate(hipStream_t s, int32_t nnode, const int32_t* rowptr,
                             const int32_t* slots, const double* vm, const double* sf,
                             double* acc, int first) {
  if (nnode <= 0) return;
  hipLaunchKernelGGL(k_nodal_accumulate, dim3(nblk(nnode, 256)), dim3(256), 0, s, nnode,
                     rowptr, slots, vm, sf, acc, first);
}

void launch_nodal_acc_out(hipStream_t s, int32_t nnode, const int32_t* rowptr,
                          const double* acc, float* vmo, float* scfo) {
  if (nnode <= 0) return;
  hipLaunchKernelGGL(k_nodal_acc_out, dim3(nblk(nnode, 256)), dim3(256), 0, s, nnode,
                     rowptr, acc, vmo, scfo);
}

void launch_nodal_fields(hipStream_t s, int32_t nnode, const int32_t* rowptr,
                         const int32_t* slots, const double* stress, float* sig,
                         float* psig, float* nsig) {
  if (nnode <= 0 || !(sig || psig || nsig)) return;
  hipLaunchKernelGGL(k_nodal_fields, dim3(nblk(nnode, 256)), dim3(256), 0, s, nnode,
                     rowptr, slots, stress, sig, psig, nsig);
}

void launch_excav_load(hipStream_t s, int32_t nelmt, const double* corners,
                       const int32_t* mat_id, const double* gam, const double* stress,
                       double* Y) {
  if (nelmt <= 0) return;
  hipLaunchKernelGGL(k_excav_load, dim3(nelmt), dim3(128), 0, s, nelmt, corners,
                     mat_id, gam, stress, Y);
}

void launch_plastic_body(hipStream_t s, const BodyArgs& b) {
  if (b.nel <= 0 || b.ncols <= 0) return;
  hipLaunchKernelGGL(k_plastic_body, dim3(nblk(b.nel, BODY_EPB), b.ncols),
                     dim3(BODY_THREADS), 0, s, b.nel, b.elmt, b.corners, b.gdof,
                     b.mat_id, b.mp, b.nmat, b.x, b.xs, b.evpt, b.stress, b.BL,
                     b.flags, b.ps, b.is_last, b.is_last_col, b.cols, b.excav,
                     b.vm_pt, b.sf_pt);
}

// The device's constitutive point functions on given records (stress(6),
// phi, c, psi): sigm, dsbar, theta (stress_invariant), f (mohcouf), dq1..3
// (mohcouq) -- the values the plastic body computes at a GLL point.
__global__ void k_constitutive(int32_t n, const double* __restrict__ rec,
                               double* __restrict__ out) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* r = rec + (int64_t)i * 9;
  double sigm, dsbar, theta, dq1, dq2, dq3;
  stress_invariant(r, sigm, dsbar, theta);
  const double f = mohcouf(r[6], r[7], sigm, dsbar, theta);
  mohcouq(r[8], dsbar, theta, dq1, dq2, dq3);
  double* o = out + (int64_t)i * 7;
  o[0] = sigm;
  o[1] = dsbar;
  o[2] = theta;
  o[3] = f;
  o[4] = dq1;
  o[5] = dq2;
  o[6] = dq3;
}

// The device math library's asin, sin, cos or tan (fn 0..3) of x.
__global__ void k_math(int32_t n, int fn, const double* __restrict__ x,
                       double* __restrict__ y) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = x[i];
  y[i] = fn == 0 ? asin(v) : fn == 1 ? sin(v) : fn == 2 ? cos(v) : tan(v);
}

void launch_constitutive(hipStream_t s, int32_t n, const double* rec, double* out) {
  hipLaunchKernelGGL(k_constitutive, dim3((n + 255) / 256), dim3(256), 0, s, n, rec, out);
}

void launch_math(hipStream_t s, int32_t n, int fn, const double* x, double* y) {
  hipLaunchKernelGGL(k_math, dim3((n + 255) / 256), dim3(256), 0, s, n, fn, x, y);
}

}  // namespace sgeo
