// Internal interface between the host orchestration (sgeo_capi.cpp) and the
// kernels (sgeo_kernels.hip).  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sgeo_device.h"
#include "sgeo_element.h"

namespace sgeo {

// Scalars of one plastic iteration (src/slope.F90:286-309), device resident.
struct PlasticScalars {
  double uerr;
  double xmax;
  unsigned long long fmax_bits;  // max(0, f) as ordered bits (f >= 0)
  int32_t nl_isconv;
  int32_t n_plastic;
};

// Arguments of the (single- or multi-trial) plastic body launch.
struct BodyArgs {
  int32_t nel = 0;                 // elements processed
  const int32_t* elmt = nullptr;   // element ids (NULL: 0..nel-1)
  const double* corners = nullptr;
  const int32_t* gdof = nullptr;   // gdof of the processed elements
  const int32_t* mat_id = nullptr;
  const MatParams* mp = nullptr;   // [col][nmat]
  int nmat = 0;
  const double* x = nullptr;       // x[g*xs + col]
  int xs = 1;
  double* evpt = nullptr;          // [col][nel][27][6]
  double* stress = nullptr;
  double* BL = nullptr;            // [col][nel][81]
  uint8_t* flags = nullptr;        // [col][nel][27]
  PlasticScalars* ps = nullptr;    // [col]
  int is_last = 0;                 // used when is_last_col == NULL
  const int32_t* is_last_col = nullptr;
  const int32_t* cols = nullptr;   // active columns (NULL: column 0)
  int ncols = 1;
  int excav = 0;
  double* vm_pt = nullptr;         // [col][nel][27] (excavation)
  double* sf_pt = nullptr;         // [col][nel][27] (excavation)
};

unsigned vec_reduce_blocks(int64_t n1);

void launch_stiffness(hipStream_t s, int32_t nelmt, const double* corners,
                      const int32_t* mat_id, const double* ym, const double* nu,
                      double* K, double* kdiag, double* eld);
void launch_round_f32(hipStream_t s, int64_t n, const double* a, float* b);
void launch_ebe_elem_f64(hipStream_t s, int32_t nelmt, const double* K,
                         const int32_t* gdof, const double* p, double* Y,
                         const PcgScalars* sc, int skip_mode);
void launch_ebe_elem_f32k(hipStream_t s, int32_t nelmt, const float* K,
                          const int32_t* gdof, const double* p, double* Y,
                          const PcgScalars* sc, int skip_mode);
void launch_ebe_packed_f64(hipStream_t s, int32_t nelmt, int n_cu, const double* KP,
                           const int32_t* gdof, const double* p, double* Y,
                           const PcgScalars* sc, int skip_mode);
void launch_ebe_packed_f32(hipStream_t s, int32_t nelmt, const float* KP,
                           const int32_t* gdof, const double* p, double* Y,
                           const PcgScalars* sc, int skip_mode);
void launch_pack_upper_f64(hipStream_t s, int32_t nelmt, const double* K,
                           double* KP);
void launch_pack_upper_f32(hipStream_t s, int32_t nelmt, const double* K,
                           float* KP);
void launch_copy(hipStream_t s, int64_t n_double2, const double* a, double* b);
void launch_assemble(hipStream_t s, int64_t n1, const int32_t* rowptr,
                     const int32_t* slots, const double* Y, double* out,
                     const PcgScalars* sc, int skip_mode);
void launch_assemble_acc(hipStream_t s, int64_t n1, const int32_t* rowptr,
                         const int32_t* slots, const double* Y, double* out,
                         int subtract, const PlasticScalars* ps);
void launch_invert_diag(hipStream_t s, int64_t n1, double* d);
void launch_set_i32(hipStream_t s, int32_t* a, int32_t v);
void launch_csr_build(hipStream_t s, int64_t ns, int32_t neq, const int32_t* gdof,
                      int32_t* rowptr, int32_t* slots, int32_t* cursor, int32_t* bad);
void launch_scale_by_mat(hipStream_t s, int32_t nelmt, const double* eld,
                         const int32_t* mat_id, const double* gam, double* Y);
void launch_pcg_begin(hipStream_t s, int64_t n1, const double* u, double* part,
                      PcgScalars* sc, double tol, int32_t maxiter);
void launch_pcg_init(hipStream_t s, int64_t n1, const double* f,
                     const double* kp, const double* dinv, double* r, double* z,
                     double* p, double* part, PcgScalars* sc, int seq);
void launch_pcg_assemble_dot(hipStream_t s, int64_t n1, const int32_t* rowptr,
                             const int32_t* slots, const double* Y,
                             const double* p, double* kp, double* part,
                             PcgScalars* sc, int seq);
void launch_pcg_update(hipStream_t s, int64_t n1, const double* p,
                       const double* kp, const double* dinv, double* u,
                       double* r, double* z, double* part, PcgScalars* sc,
                       int seq);
void launch_pcg_pupdate(hipStream_t s, int64_t n1, const double* z, double* p,
                        const PcgScalars* sc);
void launch_load(hipStream_t s, int64_t n1, const double* ext,
                 const double* body, double* load);
void launch_uerr(hipStream_t s, int64_t n1, double* x, double* oldx,
                 double* part, PlasticScalars* ps, double nl_tol);
void launch_plastic_body(hipStream_t s, const BodyArgs& b);
void launch_constitutive(hipStream_t s, int32_t n, const double* rec, double* out);
void launch_math(hipStream_t s, int32_t n, int fn, const double* x, double* y);
void launch_elastic_stress(hipStream_t s, int32_t nelmt, const double* corners,
                           const int32_t* gdof, const int32_t* mat_id, const double* ym,
                           const double* nu, const double* x, double* stress);
void launch_excav_load(hipStream_t s, int32_t nelmt, const double* corners,
                       const int32_t* mat_id, const double* gam, const double* stress,
                       double* Y);
void launch_nodal_fields(hipStream_t s, int32_t nnode, const int32_t* rowptr,
                         const int32_t* slots, const double* stress, float* sig,
                         float* psig, float* nsig);
void launch_nodal_accumulate(hipStream_t s, int32_t nnode, const int32_t* rowptr,
                             const int32_t* slots, const double* vm, const double* sf,
                             double* acc, int first);
void launch_nodal_acc_out(hipStream_t s, int32_t nnode, const int32_t* rowptr,
                          const double* acc, float* vmo, float* scfo);

// ---- multi-trial mixed-precision engine (sgeo_batch.hip) -----------------
// Per-column states of the mixed-precision PCG.  The inner iteration runs in
// FP32 (operator, r, p, K p) with FP64 dot products and an FP64 solution x;
// the FP64 true residual f - K x (FP64 operator) starts each solve and
// replaces the recursive FP32 residual, keeping the search direction, once
// r.z has dropped by inner_tol since the last replacement.  FP64 residuals
// are formed on the residual steps only: every period-th step of the
// engine's own step count.  A starting column waits for the next residual
// step; a column due for a replacement keeps iterating and is replaced at
// the next one.  Since a column starts on a residual step and then advances
// one iteration per step, its residual steps fall at its own iteration
// counts 0, period, 2 period, ...: everything a column computes depends
// only on its own iterations and the period, never on the other columns or
// on when the host synchronises.
enum BColState {
  BCOL_OFF = 0,      // column not in this solve
  BCOL_START = 1,    // next: FP64 residual, p = D r
  BCOL_INNER = 2,    // FP32 CG iteration
  BCOL_REPLACE = 3,  // now: FP64 residual replaces r, p = D r + beta p
  BCOL_DONE = 4,     // stop test passed
  BCOL_FAIL = 5,     // maxiter reached
};
#define BCOL_BIT(s) (1 << (s))
enum BFinWhat { BFIN_RESID = 0, BFIN_ALPHA = 1, BFIN_INNER = 2 };

// x is written on the iterations i with i % xlen == 0 only (xlen 1, 2, 4 or
// 8); such an iteration applies the deferred updates in order,
//   x_i = (...((x_{i-L} + alpha_{i-L+1} p_{i-L+1}) + ...) + alpha_i p_i,
// the same FP64 operations in the same order as one update per iteration.
// The search direction of iteration i lives in p buffer i & pmask (pmask + 1
// buffers, at least xlen of them).
struct BCol {  // per-column scalars, device resident
  double rz;        // r.z of the current residual
  double rz0;       // r.z at the last FP64 residual
  double alpha, beta, maxp;
  double tol, inner_tol;
  double rz_prev;        // r.z of the previous iteration (beta after a replacement)
  double alpha_hist[8];  // alpha of iteration k at [k & 7]
  double maxx;           // max|x| of the last x written (exact)
  double xsum;           // sum of |alpha_k| max|p_k| over the iterations not in x
  int32_t it, maxiter;  // CG iterations of this solve, limit
  int32_t pending;      // 1: p = D r + beta p before the next product
  int32_t nrep;         // FP64 residuals of this solve
  int32_t want_rep;     // 1: replace r at the next residual step
  int32_t xlen;         // x written on iterations i % xlen == 0
  int32_t npend;        // iterations not yet in x (0: x is current)
  int32_t pmask;        // p buffer of iteration i: i & pmask
  int32_t xerr;         // 1: a step's x write and the iteration count disagreed
};

// GLL numbering of a hex8 mesh, built on the device (sgeo_ingest.hip)
struct SpecDev {
  int32_t nelmt = 0, nnode = 0, neq = 0;
  int32_t* g_num = nullptr;      // [e][27] node ids, 0-based
  int32_t* gdof = nullptr;       // [node][3] equations, 1-based, 0 = fixed
  int32_t* gdof_elmt = nullptr;  // [e][81]
  double* g_coord = nullptr;     // [node][3]
  double* corners = nullptr;     // [e][8][3], EXODUS corner order
};
hipError_t hex2spec_device(hipStream_t s, int32_t nelmt, const double* d_coord8,
                           const int32_t* d_conn, int32_t nbc, const int32_t* d_bc,
                           double tol, SpecDev* out);
void spec_device_free(SpecDev* d);

// launch geometry of the persistent MFMA product (result bits do not
// depend on it; chosen by sgeo_batch_autotune)
struct BmvGeom {
  int blocks_per_cu = 3;
  int xcd_split = 1;
};

int64_t batch_partials(int64_t n1);          // per column, equation-chunk partials
int64_t batch_element_partials(int32_t nelmt);  // per column, element p.Kp shares
int64_t batch_equation_chunks(int64_t n1);
// eorder: the walk over the elements (a permutation of 0 .. nelmt - 1)
void launch_bmv_mfma_f32(hipStream_t s, int32_t nelmt, int B, int n_cu, BmvGeom geom,
                         const float* KP, const int32_t* gdof, const int32_t* eorder,
                         const float* P,
                         int64_t nv, float* Y, double* edot, const int32_t* cst,
                         const BCol* bc, int want_mask);
void launch_bmv_f64(hipStream_t s, int32_t nelmt, int B, int n_cu, const double* KP,
                    const int32_t* gdof, const double* X, double* Y,
                    const int32_t* rlist);
void launch_bupdate(hipStream_t s, int64_t n1, int B, const int32_t* rowptr,
                    const int32_t* slots, const float* Y, const float* dinv, float* r,
                    double* part, const int32_t* cst, const BCol* bc);
void launch_bxfold(hipStream_t s, int64_t n1, int B, double* x, const float* p,
                   double* part, const int32_t* cst, const BCol* bc);
void launch_bpupdate(hipStream_t s, int64_t n1, int B, const float* r,
                     const float* dinv, float* p, double* part, const int32_t* cst,
                     const BCol* bc);
void launch_bmaterialize(hipStream_t s, int64_t n1, int B, double* x, const float* p,
                         const int32_t* cst, BCol* bc);
void launch_bpromote(hipStream_t s, int B, int32_t* cst, BCol* bc, int32_t* rlist,
                     int32_t* rslot);
void launch_bresid(hipStream_t s, int64_t n1, int B, const int32_t* rowptr,
                   const int32_t* slots, const double* Y64, const double* load,
                   const float* dinv, const double* x, float* r, float* p,
                   double* part, const int32_t* cst, const int32_t* rslot);
// xw (INNER only): this step's update wrote x (checked against each column)
void launch_bfinalize(hipStream_t s, int64_t nch, int B, double* part,
                      int32_t* cst, BCol* bc, int what, int xw = 0);
void launch_bload(hipStream_t s, int64_t n1, int B, const double* ext,
                  const double* body, double* load, const int32_t* active);
void launch_buerr(hipStream_t s, int64_t n1, int B, double* x, double* oldx,
                  double* part, PlasticScalars* ps, const int32_t* active,
                  double nl_tol);
void launch_col_io(hipStream_t s, int64_t n1, int B, int j, double* a,
                   const double* src, double* dst);
void launch_to_f32(hipStream_t s, int64_t n, const double* a, float* b);
void launch_bbody_assemble(hipStream_t s, int64_t n1, int B, int32_t nel,
                           const int32_t* rowptr, const int32_t* slots,
                           const double* BL, double* body, const int32_t* active,
                           const int32_t* is_last, const PlasticScalars* ps);

}  // namespace sgeo
