// Device-side helpers shared by the sgeo kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SGEO_NGLL 27
#define SGEO_NEDOF 81
#define SGEO_NST 6
#define SGEO_KSIZE (SGEO_NEDOF * SGEO_NEDOF)
// packed upper triangle (3321 entries) padded to a multiple of 64
#define SGEO_KPACK 3328

// Threads per block of the vector (per-equation) kernels.  The grid of every
// reduction is a fixed function of the vector length, so partial sums are
// combined in the same order on every run: results are deterministic.
#define SGEO_VEC_THREADS 256
#define SGEO_VEC_ITEMS 4  // equations per thread per reduction kernel

// Scalars of one PCG solve, resident in device memory so the iteration
// loop never has to round-trip through the host.
struct PcgScalars {
  double rz;       // r.z of the current iteration (src/solver.f90:61)
  double pkp;      // p.Kp
  double alpha;
  double beta;
  double maxp;     // max|p| of the current iteration
  double maxu;     // max|u| after the u update
  double tol;
  int32_t iter;        // iterations started
  int32_t maxiter;
  int32_t done;        // 1 = converged (stop test passed), 2 = failed
  int32_t conv_iter;   // cg_iter reported to the caller
  int32_t u_nonzero;   // max|u| > 0 at entry (src/solver.f90:39)
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ __forceinline__ float wave_sumf(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide (SGEO_VEC_THREADS) sum / max with a fixed combination order.
template <int NW>
__device__ __forceinline__ double block_sum(double v, double* lds) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) lds[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NW; i++) s += lds[i];
  }
  __syncthreads();
  return s;  // valid in thread 0
}

template <int NW>
__device__ __forceinline__ double block_max(double v, double* lds) {
  v = wave_max(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) lds[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NW; i++) s = fmax(s, lds[i]);
  }
  __syncthreads();
  return s;  // valid in thread 0
}
