// sgeo_ingest.hip -- GLL numbering of a hex8 mesh on the device (the
// reference's pre-processing, src/mesh_spec.f90:17-386 and
// src/apply_bc.f90:36-137), so that a request's mesh is numbered in
// milliseconds instead of seconds of host sorting.
//
//  1. GLL points of every element in local (k, j, i) order, i fastest:
//     x = sum over the 8 corners of shape_hex8 * corner x, accumulated from
//     zero in corner order (src/mesh_spec.f90:81-103).
//  2. Coincident points merged as get_global does (:146-262): for x, then y,
//     then z, the points are sorted within the current segments and a new
//     segment starts where the coordinate jumps by more than the tolerance.
//     Here each pass is a stable radix sort by coordinate followed by a
//     stable radix sort by segment (lexicographic order), a break flag per
//     sorted point and a scan.  Which points share a node depends only on
//     the sorted values, so the grouping equals the reference's.
//  3. Nodes renumbered by first appearance (get_global_indirect_addressing,
//     :357-386); a node's coordinates are those of its last point (:119-136).
//  4. Face conditions zero a node's component (src/apply_bc.f90:36-127);
//     the free components are numbered node-major, component fastest
//     (:128-137); gdof_elmt(:,e) = gdof(:, g_num(:,e)) (src/slope.F90:139-143).
#pragma clang fp contract(off)

#include <cstring>

#include <rocprim/rocprim.hpp>

#include "sgeo_internal.h"

namespace sgeo {

namespace {

__device__ __forceinline__ double shape8(int c, int n) {
  // shape_hex8(c, i, j, k) at GLL abscissae {-1, 0, 1} (src/shape_library.f90:
  // 44-62): one_eighth * (1 -+ xi) * (1 -+ eta) * (1 -+ zeta), exact values
  const int i = n % 3, j = (n / 3) % 3, k = n / 9;
  const double xi = (double)(i - 1), eta = (double)(j - 1), zeta = (double)(k - 1);
  const double xip = 1.0 + xi, xim = 1.0 - xi;
  const double etap = 1.0 + eta, etam = 1.0 - eta;
  const double zp = 1.0 + zeta, zm = 1.0 - zeta;
  switch (c) {
    case 0: return 0.125 * xim * etam * zm;
    case 1: return 0.125 * xip * etam * zm;
    case 2: return 0.125 * xip * etap * zm;
    case 3: return 0.125 * xim * etap * zm;
    case 4: return 0.125 * xim * etam * zp;
    case 5: return 0.125 * xip * etam * zp;
    case 6: return 0.125 * xip * etap * zp;
    default: return 0.125 * xim * etap * zp;
  }
}

// thread per (element, local point): coordinates of the point, and the
// element's corner coordinates (EXODUS corner order) once per element
__global__ void k_gll_points(int32_t nel, const double* __restrict__ coord8,
                             const int32_t* __restrict__ conn, double* __restrict__ px,
                             double* __restrict__ py, double* __restrict__ pz,
                             double* __restrict__ corners) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= (int64_t)nel * 27) return;
  const int64_t e = q / 27;
  const int n = (int)(q - e * 27);
  double x = 0.0, y = 0.0, z = 0.0;
  for (int c = 0; c < 8; c++) {
    const int64_t g = (int64_t)conn[e * 8 + c] - 1;
    const double s = shape8(c, n);
    x = x + s * coord8[g * 3 + 0];
    y = y + s * coord8[g * 3 + 1];
    z = z + s * coord8[g * 3 + 2];
  }
  px[q] = x;
  py[q] = y;
  pz[q] = z;
  if (n < 8) {
    const int64_t g = (int64_t)conn[e * 8 + n] - 1;
    for (int d = 0; d < 3; d++) corners[e * 24 + n * 3 + d] = coord8[g * 3 + d];
  }
}

__global__ void k_iota(int64_t n, int32_t* __restrict__ a) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = (int32_t)i;
}

__global__ void k_fill(int64_t n, int32_t* __restrict__ a, int32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = v;
}

// segment of each sorted point (key of the second, stable sort)
__global__ void k_gather(int64_t n, const int32_t* __restrict__ src,
                         const int32_t* __restrict__ idx, int32_t* __restrict__ dst) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    dst[k] = src[idx[k]];
}

// break before sorted point k: another segment, or a coordinate jump
__global__ void k_breaks(int64_t n, const int32_t* __restrict__ idx,
                         const int32_t* __restrict__ seg_sorted,
                         const double* __restrict__ p, double tol,
                         int32_t* __restrict__ brk) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x) {
    int32_t b = 1;
    if (k > 0)
      b = (seg_sorted[k] != seg_sorted[k - 1]) ||
          (fabs(p[idx[k]] - p[idx[k - 1]]) > tol);
    brk[k] = b;
  }
}

// new segment of every point: inclusive count of breaks - 1
__global__ void k_scatter_seg(int64_t n, const int32_t* __restrict__ idx,
                              const int32_t* __restrict__ incl,
                              int32_t* __restrict__ seg) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    seg[idx[k]] = incl[k] - 1;
}

// first and last point of every group (order-independent min / max)
__global__ void k_first_last(int64_t n, const int32_t* __restrict__ seg,
                             int32_t* __restrict__ first, int32_t* __restrict__ last) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    atomicMin(&first[seg[i]], (int32_t)i);
    atomicMax(&last[seg[i]], (int32_t)i);
  }
}

__global__ void k_rank(int64_t ng, const int32_t* __restrict__ order,
                       int32_t* __restrict__ rank) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < ng;
       r += (int64_t)gridDim.x * blockDim.x)
    rank[order[r]] = (int32_t)r;
}

__global__ void k_node_of_point(int64_t n, const int32_t* __restrict__ seg,
                                const int32_t* __restrict__ rank,
                                int32_t* __restrict__ g_num) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    g_num[i] = rank[seg[i]];
}

__global__ void k_node_coord(int64_t ng, const int32_t* __restrict__ last,
                             const int32_t* __restrict__ rank, const double* __restrict__ px,
                             const double* __restrict__ py, const double* __restrict__ pz,
                             double* __restrict__ g_coord) {
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < ng;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = last[g], r = rank[g];
    g_coord[r * 3 + 0] = px[i];
    g_coord[r * 3 + 1] = py[i];
    g_coord[r * 3 + 2] = pz[i];
  }
}

// face condition rows (element 1-based, face 1..6, component 0..2): the 9
// nodes of the face (src/element.f90:64-100) lose that component
__global__ void k_apply_bc(int32_t nbc, const int32_t* __restrict__ bc,
                           const int32_t* __restrict__ g_num, int32_t* __restrict__ flag) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= (int64_t)nbc * 9) return;
  const int64_t r = q / 9;
  const int m = (int)(q - r * 9);
  const int64_t e = bc[r * 3 + 0] - 1;
  const int face = bc[r * 3 + 1], comp = bc[r * 3 + 2];
  // face 1: j = 0, 2: i = 2, 3: j = 2, 4: i = 0, 5: k = 0, 6: k = 2
  const int a = m % 3, b = m / 3;
  int i, j, k;
  switch (face) {
    case 1: i = a; j = 0; k = b; break;
    case 2: i = 2; j = a; k = b; break;
    case 3: i = a; j = 2; k = b; break;
    case 4: i = 0; j = a; k = b; break;
    case 5: i = a; j = b; k = 0; break;
    default: i = a; j = b; k = 2; break;
  }
  const int n = i + 3 * j + 9 * k;
  flag[(int64_t)g_num[e * 27 + n] * 3 + comp] = 0;
}

__global__ void k_number(int64_t n, const int32_t* __restrict__ flag,
                         const int32_t* __restrict__ excl, int32_t* __restrict__ gdof) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x)
    gdof[q] = flag[q] ? excl[q] + 1 : 0;
}

__global__ void k_gdof_elmt(int64_t nel, const int32_t* __restrict__ g_num,
                            const int32_t* __restrict__ gdof, int32_t* __restrict__ ge) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nel * 81;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = q / 81;
    const int r = (int)(q - e * 81);
    ge[q] = gdof[(int64_t)g_num[e * 27 + r / 3] * 3 + r % 3];
  }
}

inline unsigned gs(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return (unsigned)(b < 65536 ? (b > 0 ? b : 1) : 65536);
}

inline unsigned bits_for(int64_t v) {  // bits that hold 0..v
  unsigned b = 1;
  while (b < 32 && (v >> b) != 0) b++;
  return b;
}

struct Buf {
  void* p = nullptr;
  ~Buf() {
    if (p) (void)hipFree(p);
  }
};

#define ICHK(x)                    \
  do {                             \
    const hipError_t e_ = (x);     \
    if (e_ != hipSuccess) return e_; \
  } while (0)

template <typename T>
hipError_t alloc(Buf& b, int64_t n) {
  return hipMalloc(&b.p, (size_t)(n > 0 ? n : 1) * sizeof(T));
}

}  // namespace

void spec_device_free(SpecDev* d) {
  void* ptrs[] = {d->g_num, d->gdof, d->gdof_elmt, d->g_coord, d->corners};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  *d = SpecDev{};
}

namespace {

hipError_t hex2spec_run(hipStream_t s, int32_t nelmt, const double* d_coord8,
                        const int32_t* d_conn, int32_t nbc, const int32_t* d_bc,
                        double tol, SpecDev& r) {
  const int64_t n = (int64_t)nelmt * 27;
  r.nelmt = nelmt;
  Buf px, py, pz, seg, segs, idx0, idx1, idx2, brk, incl, keys, tmp;
  ICHK(alloc<double>(px, n));
  ICHK(alloc<double>(py, n));
  ICHK(alloc<double>(pz, n));
  ICHK(alloc<int32_t>(seg, n));
  ICHK(alloc<int32_t>(segs, n));
  ICHK(alloc<int32_t>(idx0, n));
  ICHK(alloc<int32_t>(idx1, n));
  ICHK(alloc<int32_t>(idx2, n));
  ICHK(alloc<int32_t>(brk, n));
  ICHK(alloc<int32_t>(incl, n));
  ICHK(alloc<double>(keys, n));
  ICHK(hipMalloc((void**)&r.corners, (size_t)nelmt * 24 * sizeof(double)));
  auto* X = (double*)px.p;
  auto* Y = (double*)py.p;
  auto* Z = (double*)pz.p;
  auto* SEG = (int32_t*)seg.p;
  auto* SEGS = (int32_t*)segs.p;
  auto* I0 = (int32_t*)idx0.p;
  auto* I1 = (int32_t*)idx1.p;
  auto* I2 = (int32_t*)idx2.p;
  auto* BRK = (int32_t*)brk.p;
  auto* INC = (int32_t*)incl.p;
  auto* KEYS = (double*)keys.p;
  hipLaunchKernelGGL(k_gll_points, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     nelmt, d_coord8, d_conn, X, Y, Z, r.corners);
  ICHK(hipGetLastError());
  // temporary storage: the largest of the sorts and scans below
  size_t tb = 0, t1 = 0;
  ICHK(rocprim::radix_sort_pairs(nullptr, t1, X, KEYS, I0, I1, (size_t)n, 0, 64, s));
  tb = std::max(tb, t1);
  ICHK(rocprim::radix_sort_pairs(nullptr, t1, SEGS, BRK, I1, I2, (size_t)n, 0, 32, s));
  tb = std::max(tb, t1);
  ICHK(rocprim::inclusive_scan(nullptr, t1, BRK, INC, (size_t)n, rocprim::plus<int32_t>(),
                               s));
  tb = std::max(tb, t1);
  ICHK(rocprim::exclusive_scan(nullptr, t1, BRK, INC, 0, (size_t)n * 3,
                               rocprim::plus<int32_t>(), s));
  tb = std::max(tb, t1);
  ICHK(hipMalloc(&tmp.p, tb));
  hipLaunchKernelGGL(k_fill, dim3(gs(n)), dim3(256), 0, s, n, SEG, 0);
  int32_t nseg = 1;
  const double* P[3] = {X, Y, Z};
  for (int d = 0; d < 3; d++) {
    hipLaunchKernelGGL(k_iota, dim3(gs(n)), dim3(256), 0, s, n, I0);
    size_t t = tb;
    // by coordinate, then (stably) by segment: lexicographic (segment, value)
    ICHK(rocprim::radix_sort_pairs(tmp.p, t, P[d], KEYS, I0, I1, (size_t)n, 0, 64, s));
    hipLaunchKernelGGL(k_gather, dim3(gs(n)), dim3(256), 0, s, n, SEG, I1, SEGS);
    t = tb;
    ICHK(rocprim::radix_sort_pairs(tmp.p, t, SEGS, BRK, I1, I2, (size_t)n, 0,
                                   bits_for(nseg), s));
    // BRK now holds the sorted segments; I2 the lexicographic order
    hipLaunchKernelGGL(k_breaks, dim3(gs(n)), dim3(256), 0, s, n, I2, BRK, P[d], tol, SEGS);
    t = tb;
    ICHK(rocprim::inclusive_scan(tmp.p, t, SEGS, INC, (size_t)n, rocprim::plus<int32_t>(),
                                 s));
    hipLaunchKernelGGL(k_scatter_seg, dim3(gs(n)), dim3(256), 0, s, n, I2, INC, SEG);
    ICHK(hipMemcpyAsync(&nseg, INC + (n - 1), 4, hipMemcpyDeviceToHost, s));
    ICHK(hipStreamSynchronize(s));
  }
  const int64_t ng = nseg;
  r.nnode = nseg;
  // first-appearance renumbering and node coordinates
  int32_t* FIRST = I0;  // reuse: ng <= n
  int32_t* LAST = I1;
  hipLaunchKernelGGL(k_fill, dim3(gs(ng)), dim3(256), 0, s, ng, FIRST, INT32_MAX);
  hipLaunchKernelGGL(k_fill, dim3(gs(ng)), dim3(256), 0, s, ng, LAST, -1);
  hipLaunchKernelGGL(k_first_last, dim3(gs(n)), dim3(256), 0, s, n, SEG, FIRST, LAST);
  hipLaunchKernelGGL(k_iota, dim3(gs(ng)), dim3(256), 0, s, ng, I2);
  {
    size_t t = tb;  // groups ordered by their first point
    ICHK(rocprim::radix_sort_pairs(tmp.p, t, FIRST, SEGS, I2, INC, (size_t)ng, 0,
                                   bits_for(n), s));
  }
  int32_t* RANK = BRK;
  hipLaunchKernelGGL(k_rank, dim3(gs(ng)), dim3(256), 0, s, ng, INC, RANK);
  ICHK(hipMalloc((void**)&r.g_num, (size_t)n * 4));
  ICHK(hipMalloc((void**)&r.g_coord, (size_t)ng * 3 * 8));
  hipLaunchKernelGGL(k_node_of_point, dim3(gs(n)), dim3(256), 0, s, n, SEG, RANK, r.g_num);
  hipLaunchKernelGGL(k_node_coord, dim3(gs(ng)), dim3(256), 0, s, ng, LAST, RANK, X, Y, Z,
                     r.g_coord);
  // boundary conditions and equation numbers
  Buf flagb, exclb;
  ICHK(alloc<int32_t>(flagb, 3 * ng));
  ICHK(alloc<int32_t>(exclb, 3 * ng));
  int32_t* FLAG = (int32_t*)flagb.p;
  int32_t* EXCL = (int32_t*)exclb.p;
  hipLaunchKernelGGL(k_fill, dim3(gs(3 * ng)), dim3(256), 0, s, 3 * ng, FLAG, 1);
  if (nbc > 0)
    hipLaunchKernelGGL(k_apply_bc, dim3((unsigned)(((int64_t)nbc * 9 + 255) / 256)),
                       dim3(256), 0, s, nbc, d_bc, r.g_num, FLAG);
  {
    size_t t = tb;
    ICHK(rocprim::exclusive_scan(tmp.p, t, FLAG, EXCL, 0, (size_t)(3 * ng),
                                 rocprim::plus<int32_t>(), s));
  }
  ICHK(hipMalloc((void**)&r.gdof, (size_t)ng * 3 * 4));
  hipLaunchKernelGGL(k_number, dim3(gs(3 * ng)), dim3(256), 0, s, 3 * ng, FLAG, EXCL, r.gdof);
  int32_t lastx = 0, lastf = 0;
  ICHK(hipMemcpyAsync(&lastx, EXCL + (3 * ng - 1), 4, hipMemcpyDeviceToHost, s));
  ICHK(hipMemcpyAsync(&lastf, FLAG + (3 * ng - 1), 4, hipMemcpyDeviceToHost, s));
  ICHK(hipMalloc((void**)&r.gdof_elmt, (size_t)nelmt * 81 * 4));
  hipLaunchKernelGGL(k_gdof_elmt, dim3(gs((int64_t)nelmt * 81)), dim3(256), 0, s,
                     (int64_t)nelmt, r.g_num, r.gdof, r.gdof_elmt);
  ICHK(hipGetLastError());
  ICHK(hipStreamSynchronize(s));
  r.neq = lastx + lastf;
  return hipSuccess;
}

}  // namespace

hipError_t hex2spec_device(hipStream_t s, int32_t nelmt, const double* d_coord8,
                           const int32_t* d_conn, int32_t nbc, const int32_t* d_bc,
                           double tol, SpecDev* out) {
  SpecDev r;
  const hipError_t e = hex2spec_run(s, nelmt, d_coord8, d_conn, nbc, d_bc, tol, r);
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(s);
    spec_device_free(&r);
    return e;
  }
  *out = r;
  return hipSuccess;
}

}  // namespace sgeo
