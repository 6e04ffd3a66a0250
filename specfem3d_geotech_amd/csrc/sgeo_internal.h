// Internal interface between the host orchestration (sgeo_capi.cpp) and the
// kernels (sgeo_kernels.hip).  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sgeo_device.h"

namespace sgeo {

// Factored material block of one strength-reduction trial
// (src/plastic_library.f90:42-88) plus the trial's pseudo-time step.
struct MatParams {
  double ym, nuf, phif, cohf, psif, dt;
};

// Scalars of one plastic iteration (src/slope.F90:286-309), device resident.
struct PlasticScalars {
  double uerr;
  double xmax;
  unsigned long long fmax_bits;  // max(0, f) as ordered bits (f >= 0)
  int32_t nl_isconv;
  int32_t n_plastic;
};

unsigned vec_reduce_blocks(int64_t n1);

void upload_tables(hipStream_t s, const double* d1, const double* w);
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
void launch_ebe_packed_f64(hipStream_t s, int32_t nelmt, const double* KP,
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
void launch_plastic_body(hipStream_t s, int32_t nel, const int32_t* elmt,
                         const double* corners, const int32_t* gdof,
                         const int32_t* mat_id, const MatParams* mp,
                         const double* x, double* evpt, double* stress,
                         double* BL, uint8_t* flags, PlasticScalars* ps,
                         int is_last, int excav, const int32_t* g_num,
                         double* vm_pt);

}  // namespace sgeo
